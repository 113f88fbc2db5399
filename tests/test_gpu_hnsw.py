"""GPU HNSW search against the oracle's restatement of vecdex's CPU search on the same graph.

The GPU computes every distance in the reference's summation order and reproduces the search
order exactly, so results must be identical per query (ids and distance bits) — stricter than
the >= 99.5% identical / recall-within-0.002 bar."""

import ctypes

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import _native as nat
from hnsw_vector_search_amd import datasets
from oracle import oracle as o

pytestmark = pytest.mark.gpu

GREEDY_FP = 6133164406513967666   # tests/hnsw_parallel_test.cpp:164
TOP10_FP = 15666754926265556722   # tests/hnsw_parallel_test.cpp:165


def gpu_search(index, queries, k, ef, **kw):
    out, cnt = index.search_arrays(queries, k, ef, **kw)
    return out, cnt


def assert_identical(index, queries_raw, k, ef, metric):
    prepared = o.normalize(queries_raw) if metric == 2 else queries_raw
    ref, ref_cnt = o.hnsw_search(index.graph_arrays(), prepared, k, ef)
    got, cnt = gpu_search(index, queries_raw, k, ef)
    np.testing.assert_array_equal(cnt, ref_cnt)
    for q in range(len(queries_raw)):
        c = int(cnt[q])
        assert got[q, :c].tobytes() == ref[q, :c].tobytes(), f"query {q}, k {k}, ef {ef}"


def fingerprint(out, cnt):
    return o.fingerprint([(out[i]["id"][:cnt[i]], out[i]["distance"][:cnt[i]])
                          for i in range(len(cnt))])


def test_gpu_reproduces_the_golden_fingerprints():
    print("case: golden fingerprints, 1000 x 16", flush=True)
    base = datasets.integer_matrix(1000, 16, 1)
    queries = datasets.integer_matrix(200, 16, 2)
    index = vdx.HnswIndex(16, vdx.Metric.L2, M=8, ef_construction=64, seed=42)
    index.add(base)
    assert fingerprint(*gpu_search(index, queries, 1, 1)) == GREEDY_FP
    assert fingerprint(*gpu_search(index, queries, 10, 10)) == TOP10_FP


@pytest.fixture(scope="module")
def gmm10k():
    """10,000 x 128 integer GMM, M=16, efConstruction=200, seed 0 (the CPU-runnable config)."""
    base, queries = datasets.gmm_base_and_queries(datasets.gmm64(), 10_000, 100)
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=16, ef_construction=200, seed=0)
    index.add(base, num_threads=8)
    return index, base, queries


@pytest.mark.parametrize("ef", [10, 16, 64, 128, 512])
def test_gmm10k_identical_to_cpu_search(gmm10k, ef):
    print(f"case: gmm 10k x 128, ef {ef}", flush=True)
    index, _, queries = gmm10k
    assert_identical(index, queries, 10, ef, 0)


def test_gmm10k_recall_against_float64_brute_force(gmm10k):
    index, base, queries = gmm10k
    truth, _ = o.bruteforce_f64(base, queries, 10)
    got, _ = gpu_search(index, queries, 10, 64)
    assert o.recall_at_k(got["id"], truth, 10) >= 0.95


def test_cosine_float_data_identical():
    print("case: anisotropic cosine 20k x 100", flush=True)
    gen = datasets.AnisotropicGaussian(100, 0.8, seed=0)
    base, queries = gen.sample(20_000, 1), gen.sample(200, 2)
    index = vdx.HnswIndex(100, vdx.Metric.Cosine, M=16, ef_construction=200, seed=0)
    index.add(base, num_threads=8)
    for ef in (16, 64, 256):
        assert_identical(index, queries, 10, ef, 2)


@pytest.mark.parametrize("metric", [0, 1, 2])
@pytest.mark.parametrize("dim", [1, 3, 8, 17, 40, 64, 100, 128, 200, 287])
def test_every_metric_and_dimension(metric, dim):
    print(f"case: metric {metric} dim {dim}", flush=True)
    rng = np.random.default_rng(dim)
    base = rng.standard_normal((1500, dim), dtype=np.float32)
    queries = rng.standard_normal((40, dim), dtype=np.float32)
    index = vdx.HnswIndex(dim, metric, M=8, ef_construction=60, seed=3)
    index.add(base, num_threads=4)
    for k, ef in ((1, 1), (10, 32), (20, 5)):
        assert_identical(index, queries, k, ef, metric)


def test_exact_when_ef_covers_every_node():
    rng = np.random.default_rng(1)
    base = rng.uniform(-1, 1, (300, 8)).astype(np.float32)
    queries = rng.uniform(-1, 1, (20, 8)).astype(np.float32)
    index = vdx.HnswIndex(8, 0, M=8, ef_construction=50)
    index.add(base)
    got, _ = gpu_search(index, queries, 10, 300)
    exact = o.bruteforce_f32(base, queries, 10)
    np.testing.assert_array_equal(got[:, :10], exact)


def test_mixed_k_and_ef_in_one_launch(gmm10k):
    print("case: mixed k/ef", flush=True)
    index, _, queries = gmm10k
    rng = np.random.default_rng(7)
    ks = rng.choice([1, 10, 100], len(queries)).astype(np.uint32)
    efs = rng.choice([16, 32, 64, 128, 256, 512], len(queries)).astype(np.uint32)
    got, cnt = gpu_search(index, queries, ks, efs)
    g = index.graph_arrays()
    for q in range(len(queries)):
        ref, ref_cnt = o.hnsw_search(g, queries[q:q + 1], int(ks[q]), int(efs[q]))
        assert cnt[q] == ref_cnt[0]
        assert got[q, :cnt[q]].tobytes() == ref[0, :ref_cnt[0]].tobytes()


def test_results_are_deterministic(gmm10k):
    index, _, queries = gmm10k
    a, _ = gpu_search(index, queries, 10, 64)
    b, _ = gpu_search(index, queries, 10, 64)
    assert a.tobytes() == b.tobytes()


def test_device_entry_point_matches_host_entry_point(gmm10k):
    import ctypes
    index, _, queries = gmm10k
    host, _ = gpu_search(index, queries, 10, 64)
    handle = index.gpu_handle()
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(queries, ctx)
    out = nat.DeviceArray((len(queries), 10), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(len(queries), np.uint32, ctx)
    stats = (ctypes.c_uint64 * 8)()
    opts = nat.SearchOpts()
    opts.stats = ctypes.cast(stats, ctypes.c_void_p)
    nat.check(nat.lib.vdx_gpu_hnsw_search_device(
        handle, dq.ptr, len(queries), None, 10, None, 64, 10, out.ptr, 10, cnt.ptr,
        ctypes.byref(opts), None))
    nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
    assert out.numpy().tobytes() == host[:, :10].tobytes()
    assert stats[0] > 0 and stats[1] > stats[0]


# ---- vecdex.Index semantics on the GPU (python/tests/test_index.py) -------------------------

DIM = 16


def rv(n, seed=0):
    return np.random.default_rng(seed).standard_normal((n, DIM), dtype=np.float32)


@pytest.mark.parametrize("space", ["l2", "ip", "cosine"])
def test_knn_query_recall_and_distances(space):
    base, queries = rv(2000), rv(100, 1)
    index = vdx.Index(space, DIM, M=16, ef_construction=100)
    index.add_items(base)
    ids, dist = index.knn_query(queries, k=10, ef=64)
    b, q = base.astype(np.float64), queries.astype(np.float64)
    if space == "cosine":
        b /= np.linalg.norm(b, axis=1, keepdims=True)
        q /= np.linalg.norm(q, axis=1, keepdims=True)
    exact = ((q * q).sum(1)[:, None] - 2 * q @ b.T + (b * b).sum(1)[None]) if space == "l2" else 1 - q @ b.T
    truth = np.argsort(exact, axis=1)[:, :10]
    recall = np.mean([len(set(f) & set(t)) / 10 for f, t in zip(ids, truth)])
    assert recall >= 0.98
    np.testing.assert_allclose(dist, np.take_along_axis(exact, ids, axis=1), rtol=1e-4, atol=1e-4)
    assert ids.dtype == np.int64 and dist.dtype == np.float32


def test_padding_and_empty_index():
    base = rv(5)
    index = vdx.Index("l2", DIM)
    index.add_items(base)
    ids, dist = index.knn_query(base[:2], k=8)
    assert sorted(ids[0, :5]) == [0, 1, 2, 3, 4]
    np.testing.assert_array_equal(ids[:, 5:], -1)
    np.testing.assert_array_equal(dist[:, 5:], np.inf)
    empty = vdx.Index("cosine", DIM)
    ids, dist = empty.knn_query(rv(3), k=4)
    np.testing.assert_array_equal(ids, np.full((3, 4), -1))
    ids, dist = empty.knn_query(np.empty((0, DIM), np.float32), k=4)
    assert ids.shape == (0, 4)


def test_vector_finds_itself_and_one_d_input():
    base = rv(300)
    index = vdx.Index("l2", DIM)
    index.add_items(base)
    ids, dist = index.knn_query(base, k=1, ef=100)
    np.testing.assert_array_equal(ids[:, 0], np.arange(300))
    np.testing.assert_array_equal(dist[:, 0], 0.0)
    ids, dist = index.knn_query(base[7], k=1)
    assert ids.shape == (1, 1) and ids[0, 0] == 7


def test_search_rejects_bad_queries_like_vecdex():
    index = vdx.Index("cosine", DIM)
    index.add_items(rv(10))
    bad = rv(4)
    bad[2, 3] = np.nan
    with pytest.raises(ValueError, match="HnswIndex::search: query must not contain NaN or inf"):
        index.knn_query(bad)
    with pytest.raises(ValueError, match="HnswIndex::search_batch: query must not contain NaN"):
        index.knn_query(bad, num_threads=4)
    with pytest.raises(ValueError, match="the cosine metric can't normalize a zero query"):
        index.knn_query(np.zeros(DIM, np.float32))


def test_num_threads_does_not_change_results():
    base, queries = rv(1000), rv(50, 1)
    index = vdx.Index("l2", DIM, M=8, ef_construction=40)
    index.add_items(base)
    a = index.knn_query(queries, k=10, num_threads=1)
    for threads in (0, 3):
        b = index.knn_query(queries, k=10, num_threads=threads)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])


def test_search_after_add_sees_new_vectors():
    index = vdx.Index("l2", DIM)
    index.add_items(rv(100))
    index.knn_query(rv(1), k=1)
    extra = rv(1, 99) + 100
    index.add_items(extra)
    ids, _ = index.knn_query(extra, k=1)
    assert ids[0, 0] == 100


def _upload(arrays):
    """Uploads numpy FORMAT sections through the C ABI; returns (status, message)."""
    import ctypes
    g = arrays
    view = nat.GraphView(metric=g["metric"], dim=g["dim"], n=g["n"], M=g["M"],
                         max_level=g["max_level"], entry_point=g["entry_point"],
                         upper_blocks=g["upper_blocks"], levels=g["levels"].ctypes.data,
                         layer0=g["layer0"].ctypes.data,
                         upper=g["upper"].ctypes.data if g["upper"].size else None,
                         vectors=g["vectors"].ctypes.data)
    h = ctypes.c_void_p()
    status = nat.lib.vdx_gpu_hnsw_upload(nat.Context.get(0).handle, ctypes.byref(view), ctypes.byref(h))
    if status == 0:
        nat.lib.vdx_gpu_hnsw_free(h)
    return status, nat.lib.vdx_last_error().decode()


def test_upload_applies_check_links_rules():
    """Links vecdex's loader rejects (src/hnsw_io.cpp:472-504) are rejected at upload too: the
    kernels rely on ids existing on the block's layer and appearing once per block."""
    print("case: upload validation, 2000 x 16", flush=True)
    index = vdx.HnswIndex(16, vdx.Metric.L2, M=8, ef_construction=64, seed=42)
    index.add(datasets.integer_matrix(2000, 16, 1))
    base = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in index.graph_arrays().items()}
    assert _upload(base)[0] == 0
    b0 = 1 + 2 * base["M"]

    def mutated(fn):
        g = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in base.items()}
        fn(g)
        return _upload(g)

    node = int(np.nonzero(base["layer0"][::b0] >= 2)[0][0])
    blk = node * b0

    def dup(g):
        g["layer0"][blk + 2] = g["layer0"][blk + 1]
    status, msg = mutated(dup)
    assert status == nat.VDX_INVALID_ARGUMENT and "repeats a node" in msg

    def self_link(g):
        g["layer0"][blk + 1] = node
    status, msg = mutated(self_link)
    assert status == nat.VDX_INVALID_ARGUMENT and "self-link" in msg

    def missing(g):
        g["layer0"][blk + 1] = g["n"]
    status, msg = mutated(missing)
    assert status == nat.VDX_INVALID_ARGUMENT and "missing node" in msg

    def too_many(g):
        g["layer0"][blk] = 2 * g["M"] + 1
    status, msg = mutated(too_many)
    assert status == nat.VDX_INVALID_ARGUMENT and "links on layer 0" in msg

    # an upper-layer link to a node that does not reach that layer
    levels = base["levels"]
    top = int(np.argmax(levels))
    low = int(np.nonzero(levels == 0)[0][0])
    bu = 1 + base["M"]
    off = int(levels[:top].sum())

    def below(g):
        g["upper"][off * bu] = max(1, g["upper"][off * bu])
        g["upper"][off * bu + 1] = low
    status, msg = mutated(below)
    assert status == nat.VDX_INVALID_ARGUMENT and "below the layer" in msg


def test_large_host_batch_staged_in_groups(gmm10k):
    """A host batch of >= 4096 queries is checked and uploaded in groups: per query the same
    results as the oracle, with per-query k/ef arrays, and the expansions counted once."""
    print("case: host batch of 5,000 queries (staged groups)", flush=True)
    index, base, _ = gmm10k
    queries = datasets.gmm_base_and_queries(datasets.gmm64(), 10, 5000, query_seed=7)[1]
    rng = np.random.default_rng(3)
    ks = rng.choice([1, 10, 100], size=len(queries)).astype(np.uint32)
    efs = rng.choice([16, 40, 64, 128], size=len(queries)).astype(np.uint32)
    stats = (ctypes.c_uint64 * 8)()
    opts = nat.SearchOpts()
    opts.stats = ctypes.cast(stats, ctypes.c_void_p)
    got, cnt = index.search_arrays(queries, ks, efs, opts=opts)
    g = index.graph_arrays()
    expansions = 0
    for kk in (1, 10, 100):
        for e in (16, 40, 64, 128):
            sel = np.nonzero((ks == kk) & (efs == e))[0]
            if len(sel) == 0:
                continue
            ref, ref_cnt, taken = o.hnsw_search(g, queries[sel], kk, e, max_expansions=0)
            np.testing.assert_array_equal(cnt[sel], ref_cnt)
            assert got[sel][:, :kk].tobytes() == ref.tobytes()
            expansions += int(taken.sum())
    assert stats[0] == expansions


def test_large_host_batch_rejects_a_bad_row_in_any_group(gmm10k):
    index, _, _ = gmm10k
    queries = datasets.gmm_base_and_queries(datasets.gmm64(), 10, 6000, query_seed=8)[1]
    out = np.full((len(queries), 10), 7, dtype=nat.NEIGHBOR_DTYPE)
    cnt = np.full(len(queries), 5, dtype=np.uint32)
    for row in (10, 2047, 2048, 4500, 5999):
        bad = queries.copy()
        bad[row, 5] = np.inf
        with pytest.raises(ValueError, match="HnswIndex::search_batch: query must not contain NaN or inf"):
            nat.check(nat.lib.vdx_gpu_hnsw_search(
                index.gpu_handle(), nat.ptr(bad), len(bad), 128, None, 10, None, 40, nat.ptr(out), 10,
                nat.ptr(cnt), b"HnswIndex::search_batch", None))
        # a rejected call leaves the caller's buffers as they were
        assert (cnt == 5).all() and (out["id"] == 7).all()
    # two bad rows in different groups: the first one's error, as one sequential pass reports
    ks = np.full(len(queries), 10, dtype=np.uint32)
    for zero_k, inf_row, message in ((4100, 300, "NaN or inf"), (300, 4100, "k must be positive")):
        bad = queries.copy()
        bad[inf_row, 0] = np.nan
        kz = ks.copy()
        kz[zero_k] = 0
        with pytest.raises(ValueError, match="HnswIndex::search_batch: .*" + message):
            nat.check(nat.lib.vdx_gpu_hnsw_search(
                index.gpu_handle(), nat.ptr(bad), len(bad), 128, nat.ptr(kz), 0, None, 40, nat.ptr(out), 10,
                nat.ptr(cnt), b"HnswIndex::search_batch", None))
        assert (cnt == 5).all() and (out["id"] == 7).all()
    # the search still serves the good batch afterwards
    got, c = index.search_arrays(queries, 10, 40)
    ref, ref_cnt = o.hnsw_search(index.graph_arrays(), queries, 10, 40)
    np.testing.assert_array_equal(c, ref_cnt)
    assert got.tobytes() == ref.tobytes()


# ---- the random-gather probe (the practical ceiling of the search's reads) ------------------

def _probe_hash(x):
    x = np.uint64(x) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    return int(x)


@pytest.mark.parametrize("row_bytes", [128, 64])
def test_gather_probe_reads_the_rows_it_names(row_bytes):
    """Each wave's sink word is the XOR of every word of the rows its hash names (recomputed
    here), so the probe reads what it claims and nothing outside the table."""
    print(f"case: gather probe, row_bytes {row_bytes}", flush=True)
    rows, waves, per_wave = 1000, 8, 64
    words = row_bytes // 4
    data = np.random.default_rng(5).integers(0, 2**32, (rows, words), dtype=np.uint64).astype(np.uint32)
    ctx = nat.Context.get(0)
    table = nat.DeviceArray.from_numpy(data, ctx)
    sink = nat.DeviceArray(waves, np.uint32, ctx)
    got_rows = ctypes.c_uint64(0)
    nat.check(nat.lib.vdx_gpu_gather_probe(ctx.handle, table.ptr, rows * row_bytes, row_bytes, per_wave, waves,
                                           sink.ptr, ctypes.byref(got_rows)))
    assert got_rows.value == waves * per_wave
    row_xor = np.bitwise_xor.reduce(data, axis=1)
    groups = 64 // (row_bytes // 16)
    want = []
    for w in range(waves):
        acc = 0
        for it in range(0, per_wave, groups * 4):
            for r in range(4):
                for g in range(groups):
                    h = _probe_hash((w * 0x9E3779B9 + (it + r * groups + g) * 0x85EBCA6B + 0x632BE5AB) & 0xFFFFFFFF)
                    acc ^= int(row_xor[(h * rows) >> 32])
        want.append(acc)
    assert sink.numpy().tolist() == want
    for bad in ((24, 8), (128, 6)):
        assert nat.lib.vdx_gpu_gather_probe(ctx.handle, table.ptr, rows * row_bytes, bad[0], per_wave, bad[1],
                                            sink.ptr, None) != 0
