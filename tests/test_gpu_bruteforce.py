"""GPU brute force against the oracle's restatement of BruteForceIndex::search (bit-exact) and
against float64 (ids equal up to ties)."""

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import datasets
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def line_index(points):
    index = vdx.BruteForceIndex(1)
    for p in points:
        index.add([p])
    return index


def pairs(result):
    return [(n.distance, n.id) for n in result]


@pytest.mark.parametrize("precision", ["exact", "bf16", "fp32"])
def test_hand_cases(precision):
    # tests/brute_force_test.cpp:46-73, 158-175
    idx = line_index([10, 0, 7, 3, -2])
    assert pairs(idx.search([2.0], 3, precision)) == [(1, 3), (4, 1), (16, 4)]
    assert pairs(line_index([5, 1]).search([0.0], 10, precision)) == [(1, 1), (25, 0)]
    ties = line_index([4, -4, 4, 9, -4])
    assert pairs(ties.search([0.0], 4, precision)) == [(16, 0), (16, 1), (16, 2), (16, 4)]
    assert pairs(ties.search([0.0], 2, precision)) == [(16, 0), (16, 1)]
    ip = vdx.BruteForceIndex(2, vdx.Metric.InnerProduct)
    for v in ([1, 0], [0, 2], [3, 3]):
        ip.add(v)
    assert pairs(ip.search([1.0, 1.0], 3, precision)) == [(-5, 2), (-1, 1), (0, 0)]
    cos = vdx.BruteForceIndex(2, vdx.Metric.Cosine)
    for v in ([10, 0], [0, 0.5], [-3, 0]):
        cos.add(v)
    assert pairs(cos.search([2.0, 0.0], 3, precision)) == [(0, 0), (1, 1), (2, 2)]
    assert vdx.BruteForceIndex(3).search([1.0, 2.0, 3.0], 5, precision) == []


@pytest.mark.parametrize("metric", [0, 1, 2])
@pytest.mark.parametrize("dim", [1, 5, 24, 100, 128, 130, 287])
@pytest.mark.parametrize("k", [1, 10, 100])
def test_exact_mode_is_bit_equal_to_cpu(metric, dim, k):
    print(f"case: bf exact metric {metric} dim {dim} k {k}", flush=True)
    rng = np.random.default_rng(dim * 10 + k)
    base = rng.uniform(-1, 1, (3000, dim)).astype(np.float32)
    queries = rng.uniform(-1, 1, (37, dim)).astype(np.float32)
    index = vdx.BruteForceIndex(dim, metric)
    index.add(base)
    got, cnt = index.search_arrays(queries, k, "exact")
    prepared = o.normalize(queries) if metric == 2 else queries
    ref = o.bruteforce_f32(index.vectors(), prepared, k, metric)
    assert (cnt == min(k, 3000)).all()
    assert got.tobytes() == ref.tobytes()


def test_exact_mode_matches_float64_ids_on_integer_data():
    base, queries = datasets.gmm_base_and_queries(datasets.gmm64(), 10_000, 100)
    index = vdx.BruteForceIndex(128)
    index.add(base)
    got, _ = index.search_arrays(queries, 10, "exact")
    ids, dist = o.bruteforce_f64(base, queries, 10)
    np.testing.assert_array_equal(got["id"], ids)  # integer data: exact distances, no ties lost
    np.testing.assert_array_equal(got["distance"], dist.astype(np.float32))


def test_full_ranking_when_k_is_n():
    # tests/brute_force_test.cpp:111-151
    rng = np.random.default_rng(42)
    base = rng.uniform(-1, 1, (500, 16)).astype(np.float32)
    index = vdx.BruteForceIndex(16)
    index.add(base)
    q = rng.uniform(-1, 1, (20, 16)).astype(np.float32)
    got, _ = index.search_arrays(q, 500, "exact")
    for r in range(20):
        assert sorted(got[r]["id"].tolist()) == list(range(500))
        for k in (1, 10, 100):
            sub, _ = index.search_arrays(q[r:r + 1], k, "exact")
            assert sub[0].tobytes() == got[r, :k].tobytes()


# ---- MFMA shortlist modes: every mode returns exactly the exact scan's result -----------------

@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp8"])
@pytest.mark.parametrize("metric", [0, 1, 2])
@pytest.mark.parametrize("dim", [5, 24, 100, 128])
def test_mfma_modes_equal_exact_small_base(precision, metric, dim):
    """n = 3000: every row reaches the shortlist step (no sampling)."""
    print(f"case: bf {precision} metric {metric} dim {dim} n 3000", flush=True)
    rng = np.random.default_rng(dim + 7 * metric)
    base = rng.uniform(-1, 1, (3000, dim)).astype(np.float32)
    queries = rng.uniform(-1, 1, (40, dim)).astype(np.float32)
    index = vdx.BruteForceIndex(dim, metric)
    index.add(base)
    for k in (1, 10, 100):
        exact, _ = index.search_arrays(queries, k, "exact")
        got, cnt = index.search_arrays(queries, k, precision)
        assert got.tobytes() == exact.tobytes(), (precision, metric, dim, k)


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp8"])
@pytest.mark.parametrize("metric,dim", [(0, 128), (2, 100), (1, 128), (2, 128)])
def test_mfma_modes_equal_exact_with_sampled_threshold(precision, metric, dim):
    """n = 60,000: the threshold pass, the filter and the certification all run (at dim 128
    BF16 takes the kernel whose base tiles are loaded straight into LDS)."""
    print(f"case: bf {precision} metric {metric} dim {dim} n 60000", flush=True)
    if metric == 0:
        base, queries = datasets.gmm_base_and_queries(datasets.gmm64(), 60_000, 300)
    else:
        gen = datasets.AnisotropicGaussian(dim, 0.8, seed=dim)
        base, queries = gen.sample(60_000, 1), gen.sample(300, 2)
    index = vdx.BruteForceIndex(dim, metric)
    index.add(base)
    for k in (1, 10, 100):
        exact, _ = index.search_arrays(queries, k, "exact")
        got, cnt = index.search_arrays(queries, k, precision)
        stats = index.last_stats
        print(f"  k {k}: shortlist {stats.shortlist}, exact fallbacks {stats.exact_fallback}, "
              f"max |gemm - exact| {stats.max_abs_error:.3g}", flush=True)
        assert got.tobytes() == exact.tobytes(), (precision, metric, k)
        assert stats.shortlist >= k


def test_bf16_is_exact_on_small_integers():
    """Integer data in [0, 255] is exact in BF16, so the GEMM distances are exact."""
    base, queries = datasets.gmm_base_and_queries(datasets.gmm64(), 20_000, 100)
    index = vdx.BruteForceIndex(128)
    index.add(base)
    got, _ = index.search_arrays(queries, 10, "bf16")
    assert index.last_stats.max_abs_error == 0.0
    assert index.last_stats.exact_fallback == 0


@pytest.mark.parametrize("metric,dim", [(0, 128), (2, 100)])
def test_rows_added_after_the_upload_are_appended(metric, dim):
    """BruteForceIndex::add after a search appends to the device copy (vdx_gpu_bf_append): every
    mode then searches all rows, with results equal to the host's exact scan over them."""
    print(f"case: bf append, metric {metric}, dim {dim}", flush=True)
    gen = datasets.AnisotropicGaussian(dim, 0.8, seed=6)
    index = vdx.BruteForceIndex(dim, metric)
    index.add(gen.sample(20_000, 1))
    queries = gen.sample(64, 2)
    prepared = o.normalize(queries) if metric == 2 else queries
    index.search_arrays(queries, 10, "exact")
    index.search_arrays(queries, 10, "bf16")  # builds the BF16 operands of the first rows
    handle = index.gpu_handle().value
    for step, rows in enumerate((3_000, 40_000)):  # the second outgrows the headroom
        index.add(gen.sample(rows, 10 + step))
        ref = o.bruteforce_f32(index.vectors(), prepared, 10, metric=metric)
        for precision in ("exact", "fp32", "bf16"):
            got, cnt = index.search_arrays(queries, 10, precision)
            assert (cnt == 10).all()
            assert got.tobytes() == ref.tobytes(), (step, precision)
        assert index.gpu_handle().value == handle


@pytest.mark.parametrize("precision", ["exact", "fp32", "bf16", "fp8"])
@pytest.mark.parametrize("workload", ["gmm1024_l2", "aniso_cosine"])
def test_modes_equal_the_oracle_on_the_config_generators(precision, workload):
    """The precision modes on the brute-force workloads' own generators (the 1,024-component
    mixture under L2, the anisotropic cosine base at dimension 100), 60,000 rows, k in
    {1, 10, 100}: ids and distance bits equal the oracle's restatement of
    BruteForceIndex::search directly (src/brute_force.cpp:20-37, 68-84)."""
    print(f"case: bf {precision} on {workload}, 60000 rows", flush=True)
    if workload == "gmm1024_l2":
        mix = datasets.gmm1024()
        base, queries, metric, dim = mix.sample(60_000, 1), mix.sample(200, 2), 0, 128
    else:
        gen = datasets.AnisotropicGaussian(100, 0.8, seed=0)
        base, queries, metric, dim = gen.sample(60_000, 1), gen.sample(200, 2), 2, 100
    index = vdx.BruteForceIndex(dim, metric)
    index.add(base)
    prepared = o.normalize(queries) if metric == 2 else queries
    for k in (1, 10, 100):
        ref = o.bruteforce_f32(index.vectors(), prepared, k, metric=metric, threads=16)
        got, cnt = index.search_arrays(queries, k, precision)
        assert (cnt == k).all()
        assert got.tobytes() == ref.tobytes(), (precision, workload, k)


@pytest.mark.parametrize("precision", ["bf16", "fp8"])
def test_uncertified_queries_are_widened_then_exact(precision):
    """Queries the first shortlist does not certify are redone with an 8x wider one, and only
    the ones that still fail go to the exact scan; every result equals the exact scan's."""
    print(f"case: widened shortlist, {precision}, 60,000 x 100 cosine", flush=True)
    gen = datasets.AnisotropicGaussian(100, 0.8, seed=11)
    base, queries = gen.sample(60_000, 1), gen.sample(400, 2)
    index = vdx.BruteForceIndex(100, 2)
    index.add(base)
    for k in (1, 10):
        exact, _ = index.search_arrays(queries, k, "exact")
        got, _ = index.search_arrays(queries, k, precision)
        st = index.last_stats
        print(f"  k {k}: widened {st.widened}, exact fallbacks {st.exact_fallback}", flush=True)
        assert got.tobytes() == exact.tobytes(), (precision, k)
        assert st.exact_fallback <= st.widened <= len(queries)
        if precision == "fp8":
            assert st.widened > 0  # FP8's bound rarely certifies the first shortlist here


@pytest.mark.parametrize("metric,dim", [(0, 128), (2, 100), (1, 128)])
def test_two_query_blocks_per_wave_gemm_equals_exact(metric, dim):
    """The BF16 filter's second GEMM shape (two 32-query blocks per wave, 4-wave workgroups)
    returns the exact scan's bytes; the tuner keeps either shape and still does."""
    print(f"case: BF16 GEMM shape 1, metric {metric}, dim {dim}", flush=True)
    if metric == 0:
        base, queries = datasets.gmm_base_and_queries(datasets.gmm1024(), 60_000, 700)
    else:
        gen = datasets.AnisotropicGaussian(dim, 0.8, seed=dim + 1)
        base, queries = gen.sample(60_000, 1), gen.sample(700, 2)
    index = vdx.BruteForceIndex(dim, metric)
    index.add(base)
    for k in (1, 10, 100):
        exact, _ = index.search_arrays(queries, k, "exact")
        for shape in (1, 0):
            index.set_gemm_shape(shape)
            got, _ = index.search_arrays(queries, k, "bf16")
            assert got.tobytes() == exact.tobytes(), (metric, k, shape)
    prepared = o.normalize(queries) if metric == 2 else queries
    t = index.tune(prepared, 10, reps=1)
    assert t["shape"] in (0, 1) and min(t["ms"]) > 0
    got, _ = index.search_arrays(queries, 10, "bf16")
    exact, _ = index.search_arrays(queries, 10, "exact")
    assert got.tobytes() == exact.tobytes()
