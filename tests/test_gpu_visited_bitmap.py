"""The exact visited bitmap in HBM (vdx_search_opts.visited_form = 2) of the LDS beam kernel
(R capacity above 512): one bitmap row per resident wave, cleared per query, test-and-set by
atomics. The search is vecdex's with an exact VisitedList (src/search_scratch.hpp:21-47), so
results equal the oracle's byte for byte and no distance is computed twice (no duplicates);
through the host entry point, the device entry point and a tuned or fixed uneven batch."""

import ctypes

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import _native as nat
from hnsw_vector_search_amd import datasets
from oracle import oracle as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cosine():
    print("case: 30,000 x 100 cosine, M=16", flush=True)
    gen = datasets.AnisotropicGaussian(100, 0.8, seed=3)
    base = gen.sample(30_000, 1)
    index = vdx.HnswIndex(100, vdx.Metric.Cosine, M=16, ef_construction=100, seed=0)
    index.add(base, num_threads=16)
    raw = gen.sample(400, 2)
    return index, o.normalize(raw), raw


def opts_with(form):
    stats = (ctypes.c_uint64 * 8)()
    opts = nat.SearchOpts()
    opts.visited_form = form
    opts.stats = ctypes.cast(stats, ctypes.c_void_p)
    return opts, stats


@pytest.mark.parametrize("ef", [600, 1000, 2000])
def test_host_entry_bitmap_equals_oracle(cosine, ef):
    print(f"case: bitmap, host entry, ef {ef}", flush=True)
    index, queries, raw = cosine
    ref, ref_cnt = o.hnsw_search(index.graph_arrays(), queries, 10, ef)
    for form in (1, 2):
        opts, stats = opts_with(form)
        got, cnt = index.search_arrays(raw, 10, ef, opts=opts)  # the host entry point normalizes
        np.testing.assert_array_equal(cnt, ref_cnt)
        assert got.tobytes() == ref.tobytes(), f"form {form}"
        if form == 2:
            assert stats[nat.STAT_NAMES.index("duplicates")] == 0  # an exact visited set


def test_device_entry_bitmap_twice_in_a_row(cosine):
    """Two launches in a row reuse the index's bitmap (the second waits for the first)."""
    index, queries, _ = cosine
    ctx = nat.Context.get(0)
    for nq, ef in ((400, 900), (150, 640)):
        print(f"case: bitmap, device entry, {nq} queries, ef {ef}", flush=True)
        q = queries[:nq]
        dq = nat.DeviceArray.from_numpy(q, ctx)
        out = nat.DeviceArray((nq, 100), nat.NEIGHBOR_DTYPE, ctx)
        cnt = nat.DeviceArray(nq, np.uint32, ctx)
        opts, _ = opts_with(2)
        nat.check(nat.lib.vdx_gpu_hnsw_search_device(index.gpu_handle(), dq.ptr, nq, None, 100, None, ef, 100,
                                                     out.ptr, 100, cnt.ptr, ctypes.byref(opts), None))
        ref, ref_cnt = o.hnsw_search(index.graph_arrays(), q, 100, ef)
        np.testing.assert_array_equal(cnt.numpy(), ref_cnt)
        assert out.numpy().tobytes() == ref.tobytes()


@pytest.mark.parametrize("form", [0, 2])
def test_batch_classes_with_the_bitmap(cosine, form):
    """An uneven batch whose largest class runs the LDS kernel: fixed to the bitmap (form 2) or
    tuned (form 0, the tuner trying both forms for that class); graph replays included."""
    print(f"case: batch, visited_form {form}", flush=True)
    index, queries, _ = cosine
    rng = np.random.default_rng(4)
    efs = rng.choice([32, 128, 700, 1200], size=len(queries)).astype(np.uint32)
    ks = rng.choice([1, 10], size=len(queries)).astype(np.uint32)
    g = index.graph_arrays()
    want = np.zeros((len(queries), 10), dtype=nat.NEIGHBOR_DTYPE)
    wcnt = np.zeros(len(queries), dtype=np.uint32)
    for ef in np.unique(efs):
        for k in np.unique(ks):
            sel = np.nonzero((efs == ef) & (ks == k))[0]
            ref, rc = o.hnsw_search(g, queries[sel], int(k), int(ef))
            want[sel, :k] = ref
            wcnt[sel] = rc
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(queries.astype(np.float32), ctx)
    dks = nat.DeviceArray.from_numpy(ks, ctx)
    defs = nat.DeviceArray.from_numpy(efs, ctx)
    out = nat.DeviceArray((len(queries), 10), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(len(queries), np.uint32, ctx)
    opts = nat.SearchOpts()
    opts.visited_form = form
    batch = vdx.SearchBatch(index, len(queries), 10, 1200, opts)
    if form == 0:
        classes = batch.tune(dq, len(queries), out, cnt, ks=dks, efs=defs, reps=1)
        assert {c["visited_form"] for c in classes if c["queries"]} <= {1, 2}
    for graph in (False, True, True):
        out.upload(np.zeros(out.shape, dtype=nat.NEIGHBOR_DTYPE))
        batch.run(dq, len(queries), out, cnt, ks=dks, efs=defs, graph=graph)
        got, gcnt = out.numpy(), cnt.numpy()
        np.testing.assert_array_equal(gcnt, wcnt)
        for q in range(len(queries)):
            c = int(wcnt[q])
            assert got[q, :c].tobytes() == want[q, :c].tobytes(), f"query {q}, graph {graph}"


def test_noalloc_entry_rejects_the_bitmap(cosine):
    index, queries, _ = cosine
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(queries[:8].astype(np.float32), ctx)
    out = nat.DeviceArray((8, 10), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(8, np.uint32, ctx)
    opts = nat.SearchOpts()
    opts.visited_form = 2
    with pytest.raises(ValueError, match="visited_form 2"):
        nat.check(nat.lib.vdx_gpu_hnsw_search_uneven_noalloc(index.gpu_handle(), dq.ptr, 8, None, 10, None, 900,
                                                             10, 900, out.ptr, 10, cnt.ptr, ctypes.byref(opts), None))
