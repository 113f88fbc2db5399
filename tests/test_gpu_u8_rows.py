"""The 8-bit row copy of the device index (vdx_gpu_hnsw_u8_rows, vdx_search_opts.row_format).

A dimension-128 index whose stored values are all integers in [0, 255] (SIFT-like data) keeps its
rows a second time as bytes, and the plain search reads those: the bytes convert to the very
floats the FP32 rows hold, so every distance, and every result, must equal both the FP32-row
search and the oracle's restatement of vecdex's search. Any other base keeps FP32 rows only, and
an add that brings in a row outside [0, 255] drops the copy."""

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import _native as nat
from hnsw_vector_search_amd import datasets
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def f32_rows():
    opts = nat.SearchOpts()
    opts.row_format = 1
    return opts


def assert_matches_oracle(index, queries, k, ef, **kw):
    prepared = o.normalize(queries) if index.metric() == vdx.Metric.Cosine else queries
    ref, ref_cnt = o.hnsw_search(index.graph_arrays(), prepared, k, ef)
    got, cnt = index.search_arrays(queries, k, ef, **kw)
    np.testing.assert_array_equal(cnt, ref_cnt)
    for q in range(len(queries)):
        c = int(cnt[q])
        assert got[q, :c].tobytes() == ref[q, :c].tobytes(), f"query {q}, k {k}, ef {ef}"
    return got


def byte_base(rows, seed):
    return datasets.gmm_base_and_queries(datasets.gmm64(), rows, 1, base_seed=seed)[0]


@pytest.mark.parametrize("metric", [vdx.Metric.L2, vdx.Metric.InnerProduct])
def test_byte_rows_give_the_fp32_results(metric):
    print(f"case: 8-bit rows, metric {metric}, 8000 x 128", flush=True)
    base, queries = datasets.gmm_base_and_queries(datasets.gmm64(), 8000, 200)
    index = vdx.HnswIndex(128, metric, M=12, ef_construction=80, seed=3)
    index.add(base, num_threads=8)
    assert index.gpu_u8_rows() == (metric == vdx.Metric.L2)  # the byte-row search is L2's
    # fractional queries too: the rows convert exactly, whatever the query holds
    frac = queries + np.float32(0.37)
    for q in (queries, frac):
        for k, ef in ((1, 1), (10, 36), (100, 300), (10, 600)):
            got = assert_matches_oracle(index, q, k, ef)
            ref, _ = index.search_arrays(q, k, ef, opts=f32_rows())
            assert got.tobytes() == ref.tobytes()


@pytest.mark.parametrize("case", ["fractional", "negative", "above 255", "dimension 100", "cosine",
                                  "inner product"])
def test_other_bases_keep_fp32_rows_only(case):
    print(f"case: no 8-bit rows, {case}", flush=True)
    dim, metric = 128, vdx.Metric.L2
    base = byte_base(2000, 4)
    if case == "fractional":
        base[17, 5] += 0.5
    elif case == "negative":
        base[1999, 127] = -1.0
    elif case == "above 255":
        base[0, 0] = 256.0
    elif case == "dimension 100":
        dim, base = 100, base[:, :100].copy()
    elif case == "cosine":
        metric = vdx.Metric.Cosine
    else:
        metric = vdx.Metric.InnerProduct
    index = vdx.HnswIndex(dim, metric, M=8, ef_construction=40, seed=1)
    index.add(base, num_threads=4)
    assert not index.gpu_u8_rows()
    assert_matches_oracle(index, base[:50] + np.float32(1.0), 10, 40)


def test_adds_keep_or_drop_the_byte_rows():
    print("case: 8-bit rows across adds, 3000 x 128 + 2500 + 1 fractional row", flush=True)
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=8, ef_construction=60, seed=9)
    index.add(byte_base(3000, 1), num_threads=4)
    queries = byte_base(100, 7) + np.float32(0.25)
    assert index.gpu_u8_rows()
    handle = index.gpu_handle().value
    index.add(byte_base(2500, 2), num_threads=8)  # grows the byte copy past its headroom
    assert index.gpu_handle().value == handle and index.gpu_u8_rows()
    assert_matches_oracle(index, queries, 10, 48)
    extra = byte_base(1, 3)
    extra[0, 64] = 12.5
    index.add(extra)
    assert index.gpu_handle().value == handle and not index.gpu_u8_rows()
    assert_matches_oracle(index, queries, 10, 48)


def test_device_only_load_keeps_the_byte_rows(tmp_path):
    print("case: 8-bit rows after a FORMAT v2 device load, 4000 x 128", flush=True)
    base, queries = datasets.gmm_base_and_queries(datasets.gmm64(), 4000, 100)
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=10, ef_construction=60, seed=2)
    index.add(base, num_threads=4)
    path = str(tmp_path / "u8.vdx2")
    index.save(path, version=2)
    dev = vdx.HnswDeviceIndex.load(path)
    assert dev.gpu_u8_rows()
    a, _ = dev.search_arrays(queries, 10, 64)
    b, _ = index.search_arrays(queries, 10, 64, opts=f32_rows())
    assert a.tobytes() == b.tobytes()
    assert_matches_oracle(index, queries, 10, 64)


def test_row_format_is_validated():
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=8, ef_construction=40, seed=1)
    index.add(byte_base(500, 5))
    opts = nat.SearchOpts()
    opts.row_format = 2
    with pytest.raises(ValueError, match="row_format"):
        index.search_arrays(byte_base(3, 6), 5, 10, opts=opts)


def test_captured_batch_follows_the_byte_rows_across_adds():
    print("case: batch graph over 8-bit rows across adds, 3000 x 128", flush=True)
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=8, ef_construction=60, seed=11)
    index.add(byte_base(3000, 1), num_threads=4)
    queries = byte_base(256, 7) + np.float32(0.5)
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(queries, ctx)
    out = nat.DeviceArray((len(queries), 10), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(len(queries), np.uint32, ctx)
    batch = vdx.SearchBatch(index, len(queries), 10, 64)
    extra = byte_base(1, 5)
    extra[0, 3] = 300.0  # the last add drops the byte copy
    for step, rows in enumerate((byte_base(4000, 2), extra, None)):
        for replay in range(2):
            batch.run(dq, len(queries), out, cnt, k=10, ef=64, graph=True)
            nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
            ref, ref_cnt = o.hnsw_search(index.graph_arrays(), queries, 10, 64)
            np.testing.assert_array_equal(cnt.numpy(), ref_cnt)
            assert out.numpy().tobytes() == ref.tobytes(), f"step {step}, replay {replay}"
        if rows is not None:
            index.add(rows, num_threads=8)
            index.gpu_handle()  # the update; the next replay must see the new buffers
            assert index.gpu_u8_rows() == (step == 0)


@pytest.mark.parametrize("n", [5, 64, 70])
@pytest.mark.parametrize("metric", [vdx.Metric.L2, vdx.Metric.Cosine])
def test_bf16_lds_fed_filter_on_tiny_bases(n, metric):
    """Fewer rows than one 64-row tile, exactly one, one and a bit: the LDS-fed BF16 kernel
    repeats the last row past the end and tests the row bound in the last tile."""
    print(f"case: bf16 128-d brute force, n {n}, metric {metric}", flush=True)
    rng = np.random.default_rng(n)
    base = rng.uniform(-1, 1, (n, 128)).astype(np.float32)
    queries = rng.uniform(-1, 1, (33, 128)).astype(np.float32)
    index = vdx.BruteForceIndex(128, metric)
    index.add(base)
    for k in (1, 5, n):
        exact, _ = index.search_arrays(queries, k, "exact")
        got, _ = index.search_arrays(queries, k, "bf16")
        assert got.tobytes() == exact.tobytes(), (n, k)


@pytest.mark.parametrize("case", ["all bytes", "one fraction", "one negative", "one above 255", "one -0.0"])
def test_integer_arithmetic_only_when_every_query_is_byte_valued(case):
    """A launch whose queries all have every component an integer in [0, 255] takes the exact
    integer arithmetic over the byte rows; one query outside that sends the launch to the float
    arithmetic. Either way every result equals the oracle's and the FP32-row search's, in one
    launch and in an uneven batch."""
    print(f"case: integer-vs-float byte-row kernel, {case}", flush=True)
    base, queries = datasets.gmm_base_and_queries(datasets.gmm64(), 8000, 300)
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=16, ef_construction=100, seed=4)
    index.add(base, num_threads=8)
    assert index.gpu_u8_rows()
    q = queries.copy()
    if case == "one fraction":
        q[137, 5] += 0.5
    elif case == "one negative":
        q[0, 0] = -1.0
    elif case == "one above 255":
        q[299, 127] = 256.0
    elif case == "one -0.0":
        q[10, :] = np.float32(-0.0)
    for k, ef in ((1, 16), (10, 64), (100, 200), (10, 500)):
        got = assert_matches_oracle(index, q, k, ef)
        ref, _ = index.search_arrays(q, k, ef, opts=f32_rows())
        assert got.tobytes() == ref.tobytes()
    # uneven batch: the classify kernel zeroes the flag, the check sets it
    ctx = nat.Context.get(0)
    rng = np.random.default_rng(3)
    efs = np.array([16, 64, 200, 500], dtype=np.uint32)[rng.integers(0, 4, len(q))]
    dq = nat.DeviceArray.from_numpy(q, ctx)
    defs = nat.DeviceArray.from_numpy(efs, ctx)
    out = nat.DeviceArray((len(q), 10), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(len(q), np.uint32, ctx)
    batch = vdx.SearchBatch(index, len(q), 10, 500)
    for graph in (False, True, True):
        batch.run(dq, len(q), out, cnt, k=10, efs=defs, graph=graph)
        nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
        got = out.numpy()
        for ef in np.unique(efs):
            sel = np.nonzero(efs == ef)[0]
            ref, _ = index.search_arrays(q[sel], 10, int(ef), opts=f32_rows())
            assert got[sel].tobytes() == ref.tobytes(), (case, ef, graph)
    del batch


@pytest.mark.parametrize("value", [256.0, 0.5, -1.0, 255.25])
def test_one_non_byte_query_in_a_host_batch_takes_the_float_kernel(value):
    """The host entry point decides between the integer and the float kernel while it stages
    the queries: a single component that is not an integer in [0, 255], in one query of a batch
    of many staging chunks, must send the whole batch to the float kernel (the integer kernel
    would read it as a different byte), and a batch of bytes (with -0.0 among them) to the
    integer kernel; either way every query equals the oracle."""
    print(f"case: host batch with one non-byte component {value}", flush=True)
    base, queries = datasets.gmm_base_and_queries(datasets.gmm64(), 8000, 1000)
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=12, ef_construction=80, seed=3)
    index.add(base, num_threads=8)
    assert index.gpu_u8_rows()
    q = np.array(queries, dtype=np.float32)
    q[3, 5] = np.float32(-0.0)
    assert_matches_oracle(index, q, 10, 40)
    q[617, 77] = np.float32(value)
    assert_matches_oracle(index, q, 10, 40)
