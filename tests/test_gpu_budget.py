"""The expansion budget (vdx_search_opts.max_expansions) on the GPU, against the oracle's bounded
search (vdxo_hnsw_search_batch_bounded, pinned on the CPU by tests/test_oracle_budget.py).

A query stops after `max_expansions` layer-0 link blocks and returns the closest k it holds;
every kernel that serves a query applies it: the integer and float kernels over the 8-bit rows,
the FP32 rows, the LDS-resident beam above R capacity 512, and uneven batches (class launches,
captured graphs). Results must equal the oracle's byte for byte; the counters must say how many
expansions ran (sum of min(budget, own count)) and how many queries the budget stopped."""

import ctypes

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import _native as nat
from hnsw_vector_search_amd import datasets
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def opts_with(budget, row_format=0):
    stats = (ctypes.c_uint64 * 8)()
    opts = nat.SearchOpts()
    opts.max_expansions = budget
    opts.row_format = row_format
    opts.stats = ctypes.cast(stats, ctypes.c_void_p)
    return opts, stats


@pytest.fixture(scope="module")
def gmm():
    print("case: 40,000 x 128 integer mixture, M=16", flush=True)
    mix = datasets.gmm1024()
    base, queries = mix.sample(40_000, 1), mix.sample(300, 2)
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=16, ef_construction=100, seed=0)
    index.add(base, num_threads=16)
    assert index.gpu_u8_rows()
    return index, queries


def check(index, queries, prepared, k, ef, budget, row_format=0):
    g = index.graph_arrays()
    ref, ref_cnt, taken = o.hnsw_search(g, prepared, k, ef, threads=16, max_expansions=budget)
    _, _, natural = o.hnsw_search(g, prepared, k, ef, threads=16, max_expansions=0)
    opts, stats = opts_with(budget, row_format)
    got, cnt = index.search_arrays(queries, k, ef, opts=opts)
    np.testing.assert_array_equal(cnt, ref_cnt)
    for q in range(len(queries)):
        c = int(cnt[q])
        assert got[q, :c].tobytes() == ref[q, :c].tobytes(), f"query {q}, k {k}, ef {ef}, budget {budget}"
    assert stats[nat.STAT_NAMES.index("expansions")] == int(taken.sum())
    assert stats[nat.STAT_NAMES.index("early_stops")] == int((natural > budget).sum())
    return taken, natural


@pytest.mark.parametrize("ef", [16, 40, 200, 600])
@pytest.mark.parametrize("budget", [1, 7, 30, 60])
def test_byte_rows_integer_queries(gmm, ef, budget):
    print(f"case: budget {budget}, ef {ef}, integer queries", flush=True)
    index, queries = gmm
    for k in (1, 10):
        taken, natural = check(index, queries, queries, k, ef, budget)
    if budget < np.median(natural):
        assert (taken == budget).mean() > 0.4  # the budget really bounds the searches


@pytest.mark.parametrize("row_format", [0, 1])
def test_fractional_queries_and_fp32_rows(gmm, row_format):
    """Fractional queries take the float kernel over the bytes (row_format 0) or the FP32 rows."""
    index, queries = gmm
    q = (queries + np.float32(0.25)).astype(np.float32)
    for ef, budget in ((40, 20), (100, 45), (700, 300)):
        print(f"case: fractional queries, row_format {row_format}, ef {ef}, budget {budget}", flush=True)
        check(index, q, q, 10, ef, budget, row_format)


@pytest.mark.parametrize("metric", [vdx.Metric.Cosine, vdx.Metric.InnerProduct])
def test_float_bases(metric):
    print(f"case: float base, metric {int(metric)}", flush=True)
    gen = datasets.AnisotropicGaussian(100, 0.8, seed=2)
    base, queries = gen.sample(20_000, 1), gen.sample(200, 2)
    index = vdx.HnswIndex(100, metric, M=12, ef_construction=80, seed=1)
    index.add(base, num_threads=16)
    prepared = o.normalize(queries) if metric == vdx.Metric.Cosine else queries
    for ef, budget in ((32, 10), (128, 64), (900, 400)):
        check(index, queries, prepared, 10, ef, budget)


def test_a_budget_no_query_reaches_changes_nothing(gmm):
    index, queries = gmm
    plain, plain_cnt = index.search_arrays(queries, 10, 64)
    opts, stats = opts_with(100_000)
    got, cnt = index.search_arrays(queries, 10, 64, opts=opts)
    assert got.tobytes() == plain.tobytes() and (cnt == plain_cnt).all()
    assert stats[nat.STAT_NAMES.index("early_stops")] == 0


def test_uneven_batch_with_a_budget(gmm):
    """Class launches and their captured graph apply the budget per query."""
    print("case: uneven batch with budget 25", flush=True)
    index, queries = gmm
    nq = len(queries)
    rng = np.random.default_rng(3)
    efs = np.array([16, 64, 256, 700], dtype=np.uint32)[rng.integers(0, 4, nq)]
    ks = np.array([1, 10], dtype=np.uint32)[rng.integers(0, 2, nq)]
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(queries, ctx)
    dks = nat.DeviceArray.from_numpy(ks, ctx)
    defs = nat.DeviceArray.from_numpy(efs, ctx)
    out = nat.DeviceArray((nq, 10), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(nq, np.uint32, ctx)
    opts = nat.SearchOpts()
    opts.max_expansions = 25
    batch = vdx.SearchBatch(index, nq, 10, 700, opts=opts)
    g = index.graph_arrays()
    for graph in (False, True, True):
        batch.run(dq, nq, out, cnt, ks=dks, efs=defs, graph=graph)
        nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
        got, gcnt = out.numpy(), cnt.numpy()
        for ef in np.unique(efs):
            for k in np.unique(ks):
                sel = np.nonzero((efs == ef) & (ks == k))[0]
                if len(sel) == 0:
                    continue
                ref, ref_cnt, _ = o.hnsw_search(g, queries[sel], int(k), int(ef), threads=16, max_expansions=25)
                np.testing.assert_array_equal(gcnt[sel], ref_cnt)
                for i, q in enumerate(sel):
                    c = int(ref_cnt[i])
                    assert got[q, :c].tobytes() == ref[i, :c].tobytes(), (q, ef, k, graph)
    del batch


def test_budget_with_the_learned_stop(gmm):
    """Under the learned stop the budget still bounds every query: a model that never stops gives
    the oracle's bounded search; one that always stops after 12 expansions stops at a smaller
    budget first."""
    index, queries = gmm
    from hnsw_vector_search_amd import learned
    params = learned.init_params(np.zeros(16), np.ones(16))
    params[-1] = -100.0  # never stop
    never = learned.StopModel(params)
    opts = learned.learned_opts(never, 0.5)
    opts.max_expansions = 20
    got, cnt = index.search_arrays(queries, 10, 64, opts=opts)
    ref, ref_cnt, _ = o.hnsw_search(index.graph_arrays(), queries, 10, 64, threads=16, max_expansions=20)
    np.testing.assert_array_equal(cnt, ref_cnt)
    assert got.tobytes() == ref.tobytes()
    params[-65:-1] = 0.0
    params[-1] = 100.0  # always stop once allowed
    always = learned.StopModel(params)
    stats = (ctypes.c_uint64 * 8)()
    opts = learned.learned_opts(always, 0.5, 12, stats)
    opts.max_expansions = 5
    index.search_arrays(queries, 10, 64, opts=opts)
    assert stats[0] == 5 * len(queries) and stats[4] == len(queries)


@pytest.mark.parametrize("entry", ["host", "device"])
def test_per_query_budgets(gmm, entry):
    """vdx_search_opts.max_expansions_per_query: each query's own budget (0: none), a host array
    for the host entry point and a device array for the device one; every query equals the
    oracle's bounded search at its budget."""
    print(f"case: per-query budgets through the {entry} entry point", flush=True)
    index, queries = gmm
    rng = np.random.default_rng(5)
    budgets = rng.choice([0, 3, 12, 25, 40, 90], size=len(queries)).astype(np.uint32)
    k, ef = 10, 64
    stats = (ctypes.c_uint64 * 8)()
    opts = nat.SearchOpts()
    opts.stats = ctypes.cast(stats, ctypes.c_void_p)
    opts.max_expansions = 7  # replaced by the per-query array
    if entry == "host":
        opts.max_expansions_per_query = budgets.ctypes.data
        got, cnt = index.search_arrays(queries, k, ef, opts=opts)
    else:
        ctx = nat.Context.get(0)
        dq = nat.DeviceArray.from_numpy(queries, ctx)
        db = nat.DeviceArray.from_numpy(budgets, ctx)
        dout = nat.DeviceArray((len(queries), k), nat.NEIGHBOR_DTYPE, ctx)
        dcnt = nat.DeviceArray(len(queries), np.uint32, ctx)
        opts.max_expansions_per_query = db.ptr
        nat.check(nat.lib.vdx_gpu_hnsw_search_device(index.gpu_handle(), dq.ptr, len(queries), None, k, None, ef,
                                                     k, dout.ptr, k, dcnt.ptr, ctypes.byref(opts), None))
        got, cnt = dout.numpy(), dcnt.numpy()
    g = index.graph_arrays()
    expansions = 0
    for b in np.unique(budgets):
        sel = np.nonzero(budgets == b)[0]
        ref, ref_cnt, taken = o.hnsw_search(g, queries[sel], k, ef, threads=16, max_expansions=int(b))
        np.testing.assert_array_equal(cnt[sel], ref_cnt)
        assert got[sel].tobytes() == ref.tobytes(), f"budget {b}"
        expansions += int(taken.sum())
    assert stats[nat.STAT_NAMES.index("expansions")] == expansions
