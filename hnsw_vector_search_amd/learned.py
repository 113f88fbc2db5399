"""Learned adaptive termination of the GPU HNSW beam (no vecdex counterpart).

A 16 -> 64 -> 64 -> 1 MLP (ReLU, sigmoid; BF16 with FP32 accumulation) reads 16 search-state
features after every layer-0 expansion and stops the beam when its probability exceeds a
threshold. The beam never runs past the fixed-ef search it is embedded in, so ef is the
fallback: a threshold of 1 is exactly fixed-ef search.

Pipeline (all on the GPU through the C ABI):
  collect_traces: fixed-ef searches that record the features and the current top-10 hits
    against brute-force ground truth after every expansion (vdx_gpu_hnsw_trace_device);
    label(step) = 1 when the top-10 at that step already holds as many true neighbours as the
    finished search's result ("stopping here loses nothing").
  StopModel.train: mini-batch Adam on binary cross-entropy (vdx_gpu_mlp_train).
  calibrate: the smallest-work threshold whose recall@10 against ground truth meets a target.

Features (computed in the kernel, scale-free): best and k-th distance over the layer-0 entry
distance; the expanded candidate's and the best distance over the k-th; expansions / ef',
log2(1 + expansions) / 8; result size / ef'; fraction of the last 8 expansions that changed
the top-k; the expanded candidate's rank / ef'; inserted and fresh neighbours / 2M; expansions
since the top-k last changed / 16; ef' / 512; k / 100; worst over k-th distance; metric (L2 1).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat

FEATURES = 16
HIDDEN = 64


def init_params(mean: np.ndarray, std: np.ndarray, seed: int = 0) -> np.ndarray:
    """He-initialised parameters with the given feature standardisation."""
    rng = np.random.default_rng(seed)
    inv = 1.0 / np.maximum(std, 1e-6)
    w1 = rng.standard_normal((FEATURES, HIDDEN)) * np.sqrt(2.0 / FEATURES)
    w2 = rng.standard_normal((HIDDEN, HIDDEN)) * np.sqrt(2.0 / HIDDEN)
    w3 = rng.standard_normal(HIDDEN) * np.sqrt(1.0 / HIDDEN)
    parts = [mean, inv, w1.ravel(), np.zeros(HIDDEN), w2.ravel(), np.zeros(HIDDEN), w3, [0.0]]
    params = np.concatenate([np.asarray(p, dtype=np.float64).ravel() for p in parts]).astype(np.float32)
    assert params.size == nat.MLP_PARAMS
    return params


class StopModel:
    """A stop model resident on the GPU (vdx_gpu_mlp)."""

    def __init__(self, params: np.ndarray, ctx: nat.Context | None = None):
        self.ctx = ctx or nat.Context.get(0)
        p = np.ascontiguousarray(params, dtype=np.float32)
        h = ctypes.c_void_p()
        nat.check(nat.lib.vdx_gpu_mlp_create(self.ctx.handle, p.ctypes.data, p.size, ctypes.byref(h)))
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None):
            nat.lib.vdx_gpu_mlp_free(self.handle)
            self.handle = None

    def params(self) -> np.ndarray:
        out = np.zeros(nat.MLP_PARAMS, dtype=np.float32)
        nat.check(nat.lib.vdx_gpu_mlp_params(self.handle, out.ctypes.data, out.size))
        return out

    def train(self, features: nat.DeviceArray, labels: nat.DeviceArray, epochs: int = 20,
              batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        n = features.shape[0]
        losses = np.zeros(epochs, dtype=np.float32)
        nat.check(nat.lib.vdx_gpu_mlp_train(self.handle, features.ptr, labels.ptr, n, epochs, batch,
                                            ctypes.c_float(lr), seed, losses.ctypes.data))
        return losses

    def logits(self, features: np.ndarray) -> np.ndarray:
        f = nat.DeviceArray.from_numpy(np.ascontiguousarray(features, dtype=np.float32), self.ctx)
        out = nat.DeviceArray(f.shape[0], np.float32, self.ctx)
        nat.check(nat.lib.vdx_gpu_mlp_infer(self.handle, f.ptr, f.shape[0], out.ptr))
        return out.numpy()


def collect_traces(index, queries: np.ndarray, truth_ids: np.ndarray, k: int = 10, ef: int = 128,
                   max_steps: int = 256, chunk: int = 65536):
    """Runs the traced fixed-ef search; returns (features [N,16], labels [N], steps [nq],
    final_hits [nq]) over the recorded steps of every query."""
    ctx = nat.Context.get(0)
    handle = index.gpu_handle()
    feats, labels, all_steps, all_final = [], [], [], []
    for c0 in range(0, len(queries), chunk):
        q = np.ascontiguousarray(queries[c0:c0 + chunk], dtype=np.float32)
        nq = q.shape[0]
        t = np.ascontiguousarray(truth_ids[c0:c0 + chunk, :10], dtype=np.uint32)
        dq = nat.DeviceArray.from_numpy(q, ctx)
        dt = nat.DeviceArray.from_numpy(t, ctx)
        df = nat.DeviceArray((nq, max_steps, FEATURES), np.float32, ctx)
        dh = nat.DeviceArray((nq, max_steps), np.uint8, ctx)
        ds = nat.DeviceArray(nq, np.uint32, ctx)
        dfin = nat.DeviceArray(nq, np.uint32, ctx)
        nat.check(nat.lib.vdx_gpu_hnsw_trace_device(handle, dq.ptr, nq, k, ef, dt.ptr, max_steps,
                                                    df.ptr, dh.ptr, ds.ptr, dfin.ptr))
        f, h, steps, final = df.numpy(), dh.numpy(), ds.numpy(), dfin.numpy()
        recorded = np.minimum(steps, max_steps)
        mask = np.arange(max_steps)[None, :] < recorded[:, None]
        feats.append(f[mask])
        labels.append((h >= final[:, None])[mask].astype(np.float32))
        all_steps.append(steps)
        all_final.append(final)
    return (np.concatenate(feats), np.concatenate(labels), np.concatenate(all_steps),
            np.concatenate(all_final))


def trace_hits(index, queries: np.ndarray, truth_ids: np.ndarray, k: int = 10, ef: int = 128,
               max_steps: int = 256, chunk: int = 16384):
    """The traced fixed-ef search's per-step top-10 hits: (hits uint8 [nq, max_steps], steps [nq],
    final_hits [nq]); hits[q, s] is the hits after expansion s + 1 (valid for s < steps[q])."""
    ctx = nat.Context.get(0)
    handle = index.gpu_handle()
    hits, all_steps, all_final = [], [], []
    for c0 in range(0, len(queries), chunk):
        q = np.ascontiguousarray(queries[c0:c0 + chunk], dtype=np.float32)
        nq = q.shape[0]
        t = np.ascontiguousarray(truth_ids[c0:c0 + chunk, :10], dtype=np.uint32)
        dq = nat.DeviceArray.from_numpy(q, ctx)
        dt = nat.DeviceArray.from_numpy(t, ctx)
        df = nat.DeviceArray((nq, max_steps, FEATURES), np.float32, ctx)
        dh = nat.DeviceArray((nq, max_steps), np.uint8, ctx)
        ds = nat.DeviceArray(nq, np.uint32, ctx)
        dfin = nat.DeviceArray(nq, np.uint32, ctx)
        nat.check(nat.lib.vdx_gpu_hnsw_trace_device(handle, dq.ptr, nq, k, ef, dt.ptr, max_steps,
                                                    df.ptr, dh.ptr, ds.ptr, dfin.ptr))
        hits.append(dh.numpy())
        all_steps.append(ds.numpy())
        all_final.append(dfin.numpy())
    return np.concatenate(hits), np.concatenate(all_steps), np.concatenate(all_final)


def stop_ceiling(hits: np.ndarray, steps: np.ndarray, final: np.ndarray, target: float, k: int = 10,
                 per_query: bool = False):
    """What an oracle stop could save on these traced searches: each query stopped at the first
    expansion whose top-k hits equal its final hits (the training label's event), and the fewest
    total expansions over per-query stops whose mean recall@k still reaches `target` — stopping
    a query at the first expansion reaching h hits, h chosen per query, solved exactly as a
    multiple-choice knapsack over the hits given up (dynamic programming). Returns mean
    expansions per query; with per_query, also the stops themselves ("oracle_stops",
    "budgeted_stops": expansions per query, usable as per-query expansion budgets)."""
    nq, cap = hits.shape
    recorded = np.minimum(steps, cap)
    # first[q][h]: the first expansion (1-based) with >= h hits; 0 for h = 0
    first = np.zeros((nq, k + 1), dtype=np.int64)
    for h in range(1, k + 1):
        reach = (hits >= h) & (np.arange(cap)[None, :] < recorded[:, None])
        first[:, h] = np.where(reach.any(axis=1), reach.argmax(axis=1) + 1, steps)
    level = np.minimum(final, k).astype(np.int64)
    at_final = first[np.arange(nq), level]
    out = {"fixed_expansions": round(float(steps.mean()), 2),
           "fixed_recall": round(float(final.mean() / k), 4),
           "oracle_stop_expansions": round(float(at_final.mean()), 2)}
    if per_query:
        out["oracle_stops"] = at_final.astype(np.uint32)
    need = int(np.ceil(target * k * nq - 1e-9))
    budget = int(level.sum()) - need  # hits that may be given up
    if budget < 0:
        out["budgeted_expansions"] = None
        return out
    # best[b]: the most expansions saved giving up at most b hits over the queries so far;
    # choice[i][b]: the hits the i-th query gives up in that optimum (for the per-query stops)
    best = np.zeros(budget + 1, dtype=np.int64)
    active = np.nonzero(level > 0)[0]
    choice = np.zeros((len(active), budget + 1), dtype=np.int8) if per_query else None
    for i, q in enumerate(active):
        f = int(level[q])
        nxt = best.copy()
        for d in range(1, min(f, budget) + 1):
            save = int(at_final[q] - first[q, f - d])
            cand = best[:budget + 1 - d] + save
            if per_query:
                better = cand > nxt[d:]
                choice[i, d:][better] = d
            np.maximum(nxt[d:], cand, out=nxt[d:])
        best = nxt
    out["budgeted_expansions"] = round(float((at_final.sum() - best[budget]) / nq), 2)
    out["budgeted_recall"] = round(need / (k * nq), 4)
    if per_query:
        stops = at_final.copy()
        b = budget
        for i in range(len(active) - 1, -1, -1):
            d = int(choice[i, b])
            q = active[i]
            stops[q] = first[q, int(level[q]) - d]
            b -= d
        out["budgeted_stops"] = stops.astype(np.uint32)
    return out


def train_stop_model(features: np.ndarray, labels: np.ndarray, epochs: int = 20, batch: int = 4096,
                     lr: float = 1e-3, seed: int = 0) -> tuple[StopModel, np.ndarray]:
    mean = features.mean(axis=0)
    std = features.std(axis=0)
    model = StopModel(init_params(mean, std, seed))
    ctx = model.ctx
    df = nat.DeviceArray.from_numpy(np.ascontiguousarray(features, dtype=np.float32), ctx)
    dl = nat.DeviceArray.from_numpy(np.ascontiguousarray(labels, dtype=np.float32), ctx)
    losses = model.train(df, dl, epochs, batch, lr, seed)
    return model, losses


def learned_opts(model: StopModel, threshold: float, min_steps: int = 1, stats=None,
                 every: int = 1, max_expansions: int = 0) -> nat.SearchOpts:
    """Search options of the learned stop; max_expansions adds the expansion budget (0: none)."""
    opts = nat.SearchOpts()
    opts.termination = 1
    opts.mlp = model.handle
    opts.mlp_threshold = float(threshold)
    opts.min_steps = int(min_steps)
    opts.mlp_every = int(every)
    opts.max_expansions = int(max_expansions)
    if stats is not None:
        opts.stats = ctypes.cast(stats, ctypes.c_void_p)
    return opts


def search_device(index, dq: nat.DeviceArray, nq: int, k: int, ef: int, out: nat.DeviceArray,
                  cnt: nat.DeviceArray, opts: nat.SearchOpts | None = None) -> None:
    nat.check(nat.lib.vdx_gpu_hnsw_search_device(
        index.gpu_handle(), dq.ptr, nq, None, k, None, ef, k, out.ptr, k, cnt.ptr,
        ctypes.byref(opts) if opts is not None else None, None))


def recall_of(found_ids: np.ndarray, truth_ids: np.ndarray, k: int = 10) -> float:
    f = found_ids[:, :k]
    t = truth_ids[:, :k]
    return float((t[:, :, None] == f[:, None, :]).any(axis=2).mean())


def calibrate(index, model: StopModel, queries: np.ndarray, truth_ids: np.ndarray, k: int, ef: int,
              target: float = 0.95, min_steps: int = 1,
              grid=(0.5, 0.6, 0.7, 0.8, 0.85, 0.9, 0.93, 0.95, 0.97, 0.98, 0.99, 0.995, 0.999), every: int = 1,
              max_expansions: int = 0):
    """The threshold with the fewest expansions whose recall@k meets `target` on these queries
    (under the expansion budget, if any). Returns (threshold, recall, expansions per query, table)."""
    ctx = nat.Context.get(0)
    nq = len(queries)
    dq = nat.DeviceArray.from_numpy(np.ascontiguousarray(queries, dtype=np.float32), ctx)
    out = nat.DeviceArray((nq, k), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(nq, np.uint32, ctx)
    table = []
    for t in grid:
        stats = (ctypes.c_uint64 * 8)()
        search_device(index, dq, nq, k, ef, out, cnt, learned_opts(model, t, min_steps, stats, every,
                                                                   max_expansions))
        r = recall_of(out.numpy()["id"], truth_ids, k)
        table.append((t, r, stats[0] / nq))
    ok = [row for row in table if row[1] >= target]
    best = min(ok, key=lambda row: row[2]) if ok else max(table, key=lambda row: row[1])
    return best[0], best[1], best[2], table
