"""Uneven batches (vdx_gpu_hnsw_batch_*): queries with their own (ef, k) in one call, grouped
by R-capacity class, tuned, and replayed as a captured HIP graph. Every query's result must equal
the oracle's restatement of vecdex's search at that query's (ef, k) on the same graph."""

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import _native as nat
from hnsw_vector_search_amd import datasets, learned
from oracle import oracle as o

pytestmark = pytest.mark.gpu

EFS = np.array([16, 32, 64, 100, 128, 256, 512, 700], dtype=np.uint32)
KS = np.array([1, 10, 100], dtype=np.uint32)


@pytest.fixture(scope="module")
def setup():
    print("case: batch fixture, 20,000 x 100 cosine, M=16", flush=True)
    gen = datasets.AnisotropicGaussian(100, 0.8, seed=7)
    base = gen.sample(20_000, 1)
    index = vdx.HnswIndex(100, vdx.Metric.Cosine, M=16, ef_construction=100, seed=0)
    index.add(base, num_threads=8)
    nq = 1500
    queries = o.normalize(gen.sample(nq, 2))
    rng = np.random.default_rng(3)
    efs = EFS[rng.integers(0, len(EFS), nq)]
    ks = KS[rng.integers(0, len(KS), nq)]
    return index, queries, ks, efs


def expected(index, queries, ks, efs):
    """The oracle's search per (ef, k) group, scattered back to query order."""
    g = index.graph_arrays()
    nq = len(queries)
    out = np.zeros((nq, int(ks.max())), dtype=nat.NEIGHBOR_DTYPE)
    cnt = np.zeros(nq, dtype=np.uint32)
    for ef in np.unique(efs):
        for k in np.unique(ks):
            sel = np.nonzero((efs == ef) & (ks == k))[0]
            if len(sel) == 0:
                continue
            ref, rc = o.hnsw_search(g, queries[sel], int(k), int(ef))
            out[sel, :k] = ref
            cnt[sel] = rc
    return out, cnt


def device_inputs(queries, ks, efs, max_k):
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(queries.astype(np.float32), ctx)
    dks = nat.DeviceArray.from_numpy(ks.astype(np.uint32), ctx)
    defs = nat.DeviceArray.from_numpy(efs.astype(np.uint32), ctx)
    out = nat.DeviceArray((len(queries), max_k), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(len(queries), np.uint32, ctx)
    return dq, dks, defs, out, cnt


def assert_equal_results(got, gcnt, want, wcnt):
    np.testing.assert_array_equal(gcnt, wcnt)
    for q in range(len(wcnt)):
        c = int(wcnt[q])
        assert got[q, :c].tobytes() == want[q, :c].tobytes(), f"query {q}"


def test_mixed_batch_equals_standalone_searches(setup):
    index, queries, ks, efs = setup
    want, wcnt = expected(index, queries, ks, efs)
    print("case: mixed batch, eager", flush=True)
    batch = vdx.SearchBatch(index, len(queries), int(ks.max()), int(efs.max()))
    classes = batch.classes()
    assert [c["ef_cap"] for c in classes] == [64, 128, 256, 512, 704]
    dq, dks, defs, out, cnt = device_inputs(queries, ks, efs, int(ks.max()))
    batch.run(dq, len(queries), out, cnt, ks=dks, efs=defs, graph=False)
    assert_equal_results(out.numpy(), cnt.numpy(), want, wcnt)


def test_graph_replays_match_eager(setup):
    index, queries, ks, efs = setup
    want, wcnt = expected(index, queries, ks, efs)
    batch = vdx.SearchBatch(index, len(queries), int(ks.max()), int(efs.max()))
    dq, dks, defs, out, cnt = device_inputs(queries, ks, efs, int(ks.max()))
    blank = np.zeros(out.shape, dtype=nat.NEIGHBOR_DTYPE)
    for replay in range(4):
        print(f"case: graph replay {replay}", flush=True)
        out.upload(blank)
        cnt.upload(np.zeros(len(queries), dtype=np.uint32))
        batch.run(dq, len(queries), out, cnt, ks=dks, efs=defs, graph=True)
        assert_equal_results(out.numpy(), cnt.numpy(), want, wcnt)
        # the program's own eager work between replays (the host entry point prepares its
        # queries again: normalizes them under cosine)
        got, gc = index.search_arrays(queries[:50], 10, 64)
        ref, rc = o.hnsw_search(index.graph_arrays(), o.normalize(queries[:50]), 10, 64)
        assert got.tobytes() == ref.tobytes()


def test_tuned_shapes_keep_results(setup):
    index, queries, ks, efs = setup
    want, wcnt = expected(index, queries, ks, efs)
    batch = vdx.SearchBatch(index, len(queries), int(ks.max()), int(efs.max()))
    dq, dks, defs, out, cnt = device_inputs(queries, ks, efs, int(ks.max()))
    print("case: tune", flush=True)
    classes = batch.tune(dq, len(queries), out, cnt, ks=dks, efs=defs, reps=1)
    assert sum(c["queries"] for c in classes) == len(queries)
    for c in classes:
        if c["queries"]:
            assert c["ms"] > 0 and 10 <= c["visited_log2"] <= 13 and 1 <= c["waves_per_group"] <= 4
    out.upload(np.zeros(out.shape, dtype=nat.NEIGHBOR_DTYPE))
    batch.run(dq, len(queries), out, cnt, ks=dks, efs=defs, graph=True)
    assert_equal_results(out.numpy(), cnt.numpy(), want, wcnt)


def test_uniform_batch_and_learned_mode_match_the_plain_entry_point(setup):
    index, queries, _, _ = setup
    nq = len(queries)
    params = learned.init_params(np.zeros(16, np.float32), np.ones(16, np.float32), seed=5)
    model = learned.StopModel(params)
    opts = learned.learned_opts(model, 0.6, 4)
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(queries.astype(np.float32), ctx)
    ref_out = nat.DeviceArray((nq, 10), nat.NEIGHBOR_DTYPE, ctx)
    ref_cnt = nat.DeviceArray(nq, np.uint32, ctx)
    learned.search_device(index, dq, nq, 10, 96, ref_out, ref_cnt, opts)
    print("case: learned-mode batch, uniform ef 96 k 10", flush=True)
    batch = vdx.SearchBatch(index, nq, 10, 96, opts)
    out = nat.DeviceArray((nq, 10), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(nq, np.uint32, ctx)
    batch.run(dq, nq, out, cnt, k=10, ef=96, graph=True)
    assert out.numpy().tobytes() == ref_out.numpy().tobytes()
    np.testing.assert_array_equal(cnt.numpy(), ref_cnt.numpy())


def test_rejects_calls_beyond_its_sizes(setup):
    index, queries, ks, efs = setup
    batch = vdx.SearchBatch(index, 10, 10, 64)
    dq, dks, defs, out, cnt = device_inputs(queries[:20], ks[:20], efs[:20], 10)
    with pytest.raises(ValueError, match="more queries than the batch was created for"):
        batch.run(dq, 20, out, cnt, k=10, ef=64)
    with pytest.raises(ValueError, match="k outside"):
        batch.run(dq, 5, out, cnt, k=11, ef=64)
    with pytest.raises(ValueError, match="ef above max_ef"):
        batch.run(dq, 5, out, cnt, k=10, ef=65)
