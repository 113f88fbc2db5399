"""Learned adaptive termination: traces, the BF16 MLP, training, and the learned stop inside
the beam (no vecdex counterpart; judged by recall against brute force and by equivalence with
fixed-ef search at the degenerate thresholds)."""

import ctypes

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import _native as nat
from hnsw_vector_search_amd import datasets, learned

pytestmark = pytest.mark.gpu


def bf16(x):
    """Round float32 to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32)


def numpy_logits(params, features):
    """The model's forward pass restated in numpy: BF16 inputs, weights and activations, float64
    sums (the GPU sums in FP32, so the comparison allows FP32 rounding)."""
    o = 0
    def take(n):
        nonlocal o
        v = params[o:o + n]
        o += n
        return v
    mean, inv = take(16), take(16)
    w1, b1 = bf16(take(1024)).reshape(16, 64), take(64)
    w2, b2 = bf16(take(4096)).reshape(64, 64), take(64)
    w3, b3 = bf16(take(64)), take(1)[0]
    x = bf16((features - mean) * inv).astype(np.float64)
    h1 = bf16(np.maximum(x @ w1 + b1, 0)).astype(np.float64)
    h2 = bf16(np.maximum(h1 @ w2 + b2, 0)).astype(np.float64)
    return h2 @ w3 + b3


@pytest.fixture(scope="module")
def gmm10k():
    base, queries = datasets.gmm_base_and_queries(datasets.gmm64(), 10_000, 2000)
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=16, ef_construction=200, seed=0)
    index.add(base, num_threads=8)
    bf = vdx.BruteForceIndex(128)
    bf.add(base)
    truth, _ = bf.search_arrays(queries, 10, "bf16")
    return index, queries, truth["id"].astype(np.uint32)


def test_inference_matches_numpy_bf16_forward():
    rng = np.random.default_rng(0)
    feats = rng.standard_normal((500, 16)).astype(np.float32)
    params = learned.init_params(feats.mean(0), feats.std(0), seed=3)
    params[learned.FEATURES * 2 + 1024: learned.FEATURES * 2 + 1024 + 64] = rng.standard_normal(64) * 0.1
    model = learned.StopModel(params)
    got = model.logits(feats)
    ref = numpy_logits(params, feats)
    np.testing.assert_allclose(got, ref, rtol=2e-3, atol=2e-3)


def test_trace_agrees_with_plain_search(gmm10k):
    index, queries, truth = gmm10k
    q = queries[:500]
    f, labels, steps, final = learned.collect_traces(index, q, truth[:500], k=10, ef=64, max_steps=256)
    # plain search: same expansions, same result
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(q, ctx)
    out = nat.DeviceArray((500, 10), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(500, np.uint32, ctx)
    stats = (ctypes.c_uint64 * 8)()
    opts = nat.SearchOpts()
    opts.stats = ctypes.cast(stats, ctypes.c_void_p)
    learned.search_device(index, dq, 500, 10, 64, out, cnt, opts)
    assert stats[0] == steps.sum()
    ids = out.numpy()["id"]
    hits = (truth[:500, :, None] == ids[:, None, :]).any(axis=2).sum(axis=1)
    np.testing.assert_array_equal(hits, final)
    assert f.shape == (np.minimum(steps, 256).sum(), 16) and np.isfinite(f).all()
    assert set(np.unique(labels)) <= {0.0, 1.0}
    assert 0.05 < labels.mean() < 0.95


def test_degenerate_models_bracket_fixed_ef(gmm10k):
    index, queries, truth = gmm10k
    nq = 500
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(queries[:nq], ctx)
    out = nat.DeviceArray((nq, 10), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(nq, np.uint32, ctx)
    learned.search_device(index, dq, nq, 10, 64, out, cnt)
    fixed = out.numpy()
    params = learned.init_params(np.zeros(16), np.ones(16))
    params[-1] = -100.0  # b3: never stop
    never = learned.StopModel(params)
    learned.search_device(index, dq, nq, 10, 64, out, cnt, learned.learned_opts(never, 0.5))
    assert out.numpy().tobytes() == fixed.tobytes()
    params[-65:-1] = 0.0
    params[-1] = 100.0   # always stop once allowed
    always = learned.StopModel(params)
    stats = (ctypes.c_uint64 * 8)()
    learned.search_device(index, dq, nq, 10, 64, out, cnt, learned.learned_opts(always, 0.5, 12, stats))
    assert stats[4] == nq and stats[0] == 12 * nq


def test_training_and_calibrated_learned_search(gmm10k):
    index, queries, truth = gmm10k
    train_q, eval_q = queries[:1500], queries[1500:]
    f, labels, steps, _ = learned.collect_traces(index, train_q, truth[:1500], k=10, ef=128)
    model, losses = learned.train_stop_model(f, labels, epochs=6, batch=4096, lr=1e-3)
    print("epoch losses", losses)
    assert losses[-1] < losses[0]
    logits = model.logits(f)
    acc = ((logits > 0) == (labels > 0.5)).mean()
    assert acc > max(labels.mean(), 1 - labels.mean())
    thr, rec, work, table = learned.calibrate(index, model, train_q, truth[:1500], 10, 128, 0.95)
    print("calibration", table)
    ctx = nat.Context.get(0)
    nq = len(eval_q)
    dq = nat.DeviceArray.from_numpy(eval_q, ctx)
    out = nat.DeviceArray((nq, 10), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(nq, np.uint32, ctx)
    stats = (ctypes.c_uint64 * 8)()
    learned.search_device(index, dq, nq, 10, 128, out, cnt, learned.learned_opts(model, thr, 1, stats))
    r = learned.recall_of(out.numpy()["id"], truth[1500:])
    fixed_stats = (ctypes.c_uint64 * 8)()
    opts = nat.SearchOpts()
    opts.stats = ctypes.cast(fixed_stats, ctypes.c_void_p)
    learned.search_device(index, dq, nq, 10, 128, out, cnt, opts)
    print(f"held-out: threshold {thr}, recall {r:.4f}, expansions {stats[0] / nq:.1f} "
          f"(fixed ef=128: {fixed_stats[0] / nq:.1f})")
    assert r >= 0.93
    assert stats[0] < fixed_stats[0]


# ---- the workloads' own generators: the 1,024-component mixture (L2) and the anisotropic
#      cosine base, at 40,000 rows, judged against the oracle's brute force

@pytest.fixture(scope="module", params=["gmm1024_l2", "aniso_cosine"])
def config_base(request):
    from oracle import oracle as o
    if request.param == "gmm1024_l2":
        mix = datasets.gmm1024()
        base, queries, metric, dim = mix.sample(40_000, 1), mix.sample(3000, 2), vdx.Metric.L2, 128
        stored = base
    else:
        gen = datasets.AnisotropicGaussian(100, 0.8, seed=0)
        base, metric, dim = gen.sample(40_000, 1), vdx.Metric.Cosine, 100
        queries, stored = o.normalize(gen.sample(3000, 2)), o.normalize(base)
    print(f"case: learned stop on {request.param}, 40000 rows", flush=True)
    index = vdx.HnswIndex(dim, metric, M=16, ef_construction=200, seed=0)
    index.add(base, num_threads=16)
    truth = o.bruteforce_f32(stored, queries, 10, metric=metric, threads=16)["id"].astype(np.uint32)
    return request.param, index, queries, truth


def _device_search(index, queries, k, ef, opts=None):
    ctx = nat.Context.get(0)
    nq = len(queries)
    dq = nat.DeviceArray.from_numpy(np.ascontiguousarray(queries, dtype=np.float32), ctx)
    out = nat.DeviceArray((nq, k), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(nq, np.uint32, ctx)
    learned.search_device(index, dq, nq, k, ef, out, cnt, opts)
    return out.numpy(), cnt.numpy()


def test_never_stop_model_equals_the_oracle_on_the_config_generators(config_base):
    """The learned kernel with a model that never stops is the fixed-ef search: identical to the
    oracle's restatement of vecdex's search at the cap, for k 10 and 100."""
    from oracle import oracle as o
    name, index, queries, _ = config_base
    params = learned.init_params(np.zeros(16), np.ones(16))
    params[-1] = -100.0
    never = learned.StopModel(params)
    g = index.graph_arrays()
    for k, ef in ((10, 64), (10, 128), (100, 128)):
        got, cnt = _device_search(index, queries[:500], k, ef, learned.learned_opts(never, 0.5))
        ref, ref_cnt = o.hnsw_search(g, queries[:500], k, ef, threads=16)
        np.testing.assert_array_equal(cnt, ref_cnt)
        assert got.tobytes() == ref.tobytes(), (name, k, ef)


def test_learned_stop_on_the_config_generators(config_base):
    """Traces of 2,000 training queries labelled by the oracle's brute force, a model trained on
    them, a threshold calibrated on 500 held-out queries, then 500 evaluation queries: recall
    against brute force within 0.02 of the calibration target, with fewer expansions than the
    fixed-ef search it stops early."""
    name, index, queries, truth = config_base
    train, calib, evalq = queries[:2000], queries[2000:2500], queries[2500:]
    cap = 128
    f, labels, _, _ = learned.collect_traces(index, train, truth[:2000], k=10, ef=cap)
    model, losses = learned.train_stop_model(f, labels, epochs=6, batch=4096, lr=1e-3)
    assert losses[-1] < losses[0]
    fixed_calib, _ = _device_search(index, calib, 10, cap)
    target = min(0.95, learned.recall_of(fixed_calib["id"], truth[2000:2500]) - 0.01)
    thr, rec, work, table = learned.calibrate(index, model, calib, truth[2000:2500], 10, cap, target)
    stats = (ctypes.c_uint64 * 8)()
    got, _ = _device_search(index, evalq, 10, cap, learned.learned_opts(model, thr, 1, stats))
    fixed_stats = (ctypes.c_uint64 * 8)()
    opts = nat.SearchOpts()
    opts.stats = ctypes.cast(fixed_stats, ctypes.c_void_p)
    fixed, _ = _device_search(index, evalq, 10, cap, opts)
    r = learned.recall_of(got["id"], truth[2500:])
    r_fixed = learned.recall_of(fixed["id"], truth[2500:])
    print(f"  {name}: target {target:.4f}, threshold {thr}, eval recall {r:.4f} (fixed ef {cap}: "
          f"{r_fixed:.4f}), expansions {stats[0] / 500:.1f} vs {fixed_stats[0] / 500:.1f}", flush=True)
    assert r >= target - 0.02
    assert stats[0] < fixed_stats[0]


@pytest.mark.parametrize("every", [1, 3])
def test_learned_over_byte_rows_equals_fp32_rows(gmm10k, every):
    """The learned stop over the 8-bit rows (the integer kernel for byte-valued queries, the float
    kernel for others) computes the same distances as over the FP32 rows, hence the same features
    and stop decisions: identical results and counters, for a model consulted every `every`
    expansions. (One feature, the fresh neighbours of an expansion, counts the nodes the lossy
    visited cache forgot, so both searches use one cache size.)"""
    index, queries, truth = gmm10k
    assert index.gpu_u8_rows()
    f, labels, _, _ = learned.collect_traces(index, queries[:800], truth[:800], k=10, ef=64)
    model, _ = learned.train_stop_model(f, labels, epochs=3, batch=4096, lr=1e-3)
    q_int = queries[800:1600]
    q_frac = q_int + np.float32(0.25)
    for q in (q_int, q_frac):
        ctx = nat.Context.get(0)
        dq = nat.DeviceArray.from_numpy(q, ctx)
        results = []
        for rows in (0, 1):
            stats = (ctypes.c_uint64 * 8)()
            opts = learned.learned_opts(model, 0.8, 4, stats, every)
            opts.row_format = rows
            opts.visited_log2 = 11
            out = nat.DeviceArray((len(q), 10), nat.NEIGHBOR_DTYPE, ctx)
            cnt = nat.DeviceArray(len(q), np.uint32, ctx)
            learned.search_device(index, dq, len(q), 10, 64, out, cnt, opts)
            results.append((out.numpy().tobytes(), cnt.numpy().tobytes(), list(stats)[:5]))
        assert results[0] == results[1]
        assert results[0][2][4] > 0  # the model stopped some queries
