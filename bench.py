"""Benchmark: QPS at recall@10 = 0.95 on the 1,000,000 x 128 L2 Gaussian-mixture base, one GPU:
learned termination vs fixed ef vs the host CPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-learned] [--no-cpu]

One step is one batched GPU HNSW search of the 10,000 evaluation queries. Queries, graph and
results stay in HBM for the timed region; the launches go through the library's device entry
point on its context stream, bracketed by HIP events recorded on that stream and by stream
synchronizations. (The library owns device memory too: one HIP runtime per process.)

Workloads (seeded; graph M=16, efConstruction=200, built on all host cores, cached in /tmp):
  * 1M x 128 L2: 1,024-component Gaussian mixture, integer-valued; 10k evaluation queries.
  * 1.2M x 100 cosine: anisotropic Gaussian (lambda_i ~ i^-0.8, random rotation), unit-norm;
    10k evaluation queries (learned termination only).
Ground truth: the library's BF16 brute force, whose certification makes it equal to the exact
scan (vecdex's BruteForceIndex); the 1M x 128 leg also times FP32 vs BF16 brute force.

Fixed ef: the sweep {16, 32, 64, 128, 512} and a finer {40, 48, 56}; the operating point is the
smallest ef with recall@10 >= 0.95. Learned termination: traces of 100k training queries per base
at ef=128, a stop model trained 20 epochs (Adam, lr 1e-3, batch 4096) on up to 8M trace steps,
the threshold calibrated on 10k further queries per base for recall 0.955 at ef caps 64 and 128,
then timed on the 10k evaluation queries. The headline `value` is the learned QPS when its
evaluation recall is >= 0.95, else the fixed-ef QPS.

roofline: algorithmic bytes of the headline beam launch (query, link blocks read, rows of
evaluated neighbours, results, from the kernel's counters in a counting pass with an 8K-slot
visited cache) over its measured launch time, against 8 TB/s. This base's values are integers
in [0, 255], so the device index also keeps its rows as bytes and the plain search reads those
(128 bytes per row instead of 512; the same floats, so the same results);
`headline_over_f32_rows` times the same launches over the FP32 rows. cpu_baseline: the oracle's C
restatement of vecdex's search_batch (AVX2 order) on this machine's cores, fixed-ef operating
point, vecdex's protocol (one warmup, median of 5).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP32_MFMA_TFLOPS = 157.3
BF16_MFMA_TFLOPS = 2500.0
EF_SWEEP = (16, 32, 64, 128, 512)
EF_FINE = (34, 36, 38, 40, 44, 48, 56)
TARGET_RECALL = 0.95
K = 10
CACHE = Path(os.environ.get("VDX_BENCH_CACHE", "/tmp/vdx_bench_cache"))


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads() -> int:
    return max(1, min(os.cpu_count() or 1, 16))


def build_or_load_graph(vdx, name, base, metric, threads):
    CACHE.mkdir(parents=True, exist_ok=True)
    path = CACHE / f"{name}_n{len(base)}_M16_efc200_seed0.vdx"
    if path.exists():
        t = time.perf_counter()
        index = vdx.HnswIndex.load(path)
        log(f"{name}: loaded cached graph in {time.perf_counter() - t:.1f} s")
        return index, None
    index = vdx.HnswIndex(base.shape[1], metric, M=16, ef_construction=200, seed=0)
    t = time.perf_counter()
    index.add(base, num_threads=threads)
    build_s = time.perf_counter() - t
    log(f"{name}: built graph n={len(base)} on {threads} threads in {build_s:.1f} s")
    index.save(path)
    return index, build_s


class Runner:
    """Device-resident queries and results for one query set, and timed launches."""

    def __init__(self, nat, index, queries, k=K):
        self.nat, self.index, self.k = nat, index, k
        self.ctx = nat.Context.get(0)
        self.nq = len(queries)
        self.dq = nat.DeviceArray.from_numpy(np.ascontiguousarray(queries, dtype=np.float32), self.ctx)
        self.out = nat.DeviceArray((self.nq, k), nat.NEIGHBOR_DTYPE, self.ctx)
        self.cnt = nat.DeviceArray(self.nq, np.uint32, self.ctx)

    def launch(self, ef, opts=None):
        self.nat.check(self.nat.lib.vdx_gpu_hnsw_search_device(
            self.index.gpu_handle(), self.dq.ptr, self.nq, None, self.k, None, ef, self.k,
            self.out.ptr, self.k, self.cnt.ptr, ctypes.byref(opts) if opts is not None else None, None))

    def results(self):
        return self.out.numpy()

    def timed(self, ef, steps, warmup, opts=None):
        """Milliseconds per launch over exactly `steps` launches (HIP events on the launch
        stream, stream drained before and after), and wall milliseconds per launch."""
        nat = self.nat
        for _ in range(warmup):
            self.launch(ef, opts)
        nat.check(nat.lib.vdx_gpu_synchronize(self.ctx.handle))
        t0 = time.perf_counter()
        nat.check(nat.lib.vdx_gpu_event_record(self.ctx.handle, 0))
        for _ in range(steps):
            self.launch(ef, opts)
        nat.check(nat.lib.vdx_gpu_event_record(self.ctx.handle, 1))
        nat.check(nat.lib.vdx_gpu_synchronize(self.ctx.handle))
        wall = (time.perf_counter() - t0) * 1000 / steps
        return nat.elapsed_ms(self.ctx, 0, 1) / steps, wall

    def counters(self, ef, opts=None, visited_log2=0):
        nat = self.nat
        stats = (ctypes.c_uint64 * 8)()
        o = opts if opts is not None else nat.SearchOpts()
        saved = (o.stats, o.visited_log2)
        o.stats = ctypes.cast(stats, ctypes.c_void_p)
        o.visited_log2 = visited_log2
        self.launch(ef, o)
        o.stats, o.visited_log2 = saved
        return [int(stats[i]) for i in range(5)]


def recall(found, truth):
    """Mean recall@10 by id (vecdex src/recall.cpp:11-34, through the library)."""
    from hnsw_vector_search_amd import recall as r
    return r.mean_recall_at_k(found, truth["id"].astype(np.int32), K)


def threshold_recall(found, truth):
    """Mean distance-threshold recall@10 (src/recall.cpp:36-51, through the library)."""
    from hnsw_vector_search_amd import recall as r
    return r.mean_threshold_recall_at_k(found, truth["distance"][:, K - 1], K)


def measured_traffic(mode, ef, nq, rows):
    """HBM bytes per launch of the headline kernel from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), when it covers this
    workload (rows: "u8" for the 8-bit row copy, "f32"); else None."""
    path = ROOT / "profiles" / "hnsw_beam_traffic.json"
    if not path.exists():
        return None
    for row in json.loads(path.read_text()).get("workloads", []):
        if (row.get("mode") == mode and row.get("ef") == ef and row.get("nq") == nq
                and row.get("rows", "f32") == rows):
            return row.get("traffic_bytes")
    return None


def prepared(nat, queries, cosine):
    """Queries as the device entry points take them: unit length under cosine (detail::normalize)."""
    q = np.array(queries, dtype=np.float32, order="C", copy=True)
    if cosine:
        nat.check(nat.lib.vdx_host_normalize_rows(nat.ptr(q), q.shape[0], q.shape[1]))
    return q


def ground_truth(vdx, base, metric, queries, k=K):
    bf = vdx.BruteForceIndex(base.shape[1], metric)
    bf.add(base)
    truth, _ = bf.search_arrays(queries, k, "bf16")
    return truth, bf


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--n-cos", type=int, default=1_200_000)
    ap.add_argument("--train-queries", type=int, default=100_000, help="per base")
    ap.add_argument("--max-train-steps", type=int, default=8_000_000)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="skip the uneven-batch leg")
    ap.add_argument("--serving-nq", type=int, default=65536)
    ap.add_argument("--base10m", action="store_true",
                    help="also run the 10M x 128 M=32 base (host graph build takes minutes)")
    ap.add_argument("--n10m", type=int, default=10_000_000)
    ap.add_argument("--no-gpu-build", action="store_true", help="skip the GPU-assisted build leg")
    ap.add_argument("--no-load", action="store_true", help="skip the FORMAT v1/v2 load leg")
    ap.add_argument("--host-build-10m", action="store_true",
                    help="build the 10M graph on the host alone (default: GPU-assisted)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--efs", type=str, default=",".join(map(str, EF_SWEEP + EF_FINE)))
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run the library's copy kernel once over this many MiB (counter calibration)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (HNSW search "
              "shards by query with no exchange; see the multi-GPU design notes)", file=sys.stderr)
        return 2

    import hnsw_vector_search_amd as vdx
    from hnsw_vector_search_amd import _native as nat
    from hnsw_vector_search_amd import datasets, learned

    threads = host_threads()
    t = time.perf_counter()
    mix = datasets.gmm1024()
    base = mix.sample(args.n, 1)
    queries = mix.sample(args.nq, 2)
    log(f"generated {args.n} x 128 base and {args.nq} queries in {time.perf_counter() - t:.1f} s")
    index, build_s = build_or_load_graph(vdx, "gmm1m", base, vdx.Metric.L2, threads)

    # Ground truth (and the brute-force leg): exact scan, FP32 vs BF16 MFMA shortlist.
    t = time.perf_counter()
    truth, bf = ground_truth(vdx, base, vdx.Metric.L2, queries)
    gt_s = time.perf_counter() - t
    bf_leg = bruteforce_leg(args, nat, bf, queries, vdx.Metric.L2, threads)
    del bf

    # Fixed-ef sweep.
    run = Runner(nat, index, queries)
    sweep = []
    for ef in sorted(set(map(int, args.efs.split(",")))):
        run.launch(ef)
        res = run.results()
        ms, _ = run.timed(ef, args.steps, args.warmup)
        row = {"ef": ef, "recall": round(recall(res, truth), 4),
               "recall_threshold": round(threshold_recall(res, truth), 4),
               "ms": round(ms, 4), "qps": round(args.nq / ms * 1000, 1)}
        sweep.append(row)
        log(f"fixed ef {ef}: recall {row['recall']:.4f}, {ms:.3f} ms/pass, {row['qps']:,.0f} QPS")
    gpu_build_leg = None if args.no_gpu_build else run_gpu_build(vdx, nat, base, queries, truth, threads,
                                                                  build_s, sweep)
    load_leg = None if args.no_load else run_load(vdx, index, queries)
    config_points = [s for s in sweep if s["ef"] in EF_SWEEP and s["recall"] >= TARGET_RECALL]
    fixed = min(config_points, key=lambda s: s["ef"]) if config_points else sweep[-1]
    fine_points = [s for s in sweep if s["recall"] >= TARGET_RECALL]
    fixed_best = min(fine_points, key=lambda s: s["ef"]) if fine_points else fixed

    # Learned termination.
    learned_leg = None
    headline = {"mode": "fixed-ef", "ef": fixed_best["ef"], "opts": None, "recall": fixed_best["recall"],
                "qps": fixed_best["qps"]}
    if not args.no_learned:
        learned_leg, headline = run_learned(args, vdx, nat, datasets, learned, index, base, queries,
                                            truth, mix, threads, fixed, fixed_best, headline)
    serving_leg = None if args.no_serving else run_serving(args, vdx, nat, datasets, threads)
    base10m_leg = run_10m(args, vdx, nat, datasets, threads) if args.base10m else None

    if args.copy_probe_mb > 0:
        nbytes = args.copy_probe_mb << 20
        src_buf = nat.DeviceArray(nbytes // 4, np.uint32, run.ctx)
        dst_buf = nat.DeviceArray(nbytes // 4, np.uint32, run.ctx)
        nat.check(nat.lib.vdx_gpu_copy_device(run.ctx.handle, dst_buf.ptr, src_buf.ptr, nbytes))
        log(f"copy probe: {nbytes} bytes copied by the library's copy kernel")
        del src_buf, dst_buf

    # Headline timing: exactly `steps` launches.
    ms, wall_ms = run.timed(headline["ef"], args.steps, args.warmup, headline["opts"])
    qps = args.nq / ms * 1000
    res = run.results()
    rec = recall(res, truth)
    counts = run.counters(headline["ef"], headline["opts"], visited_log2=13)
    expansions, distances, upper_blocks, dups = counts[:4]
    d, M, nq = 128, 16, args.nq
    # The plain search reads the 8-bit row copy when the index keeps one (every stored value an
    # integer in [0, 255]: the same floats, so the same results): d bytes per row, else 4d.
    rows = "u8" if index.gpu_u8_rows() and headline["mode"] == "fixed-ef" else "f32"
    row_bytes = d if rows == "u8" else 4 * d
    alg_bytes = (nq * 4 * d + upper_blocks * 4 * (1 + M) + expansions * 4 * (1 + 2 * M)
                 + (distances - dups) * row_bytes + nq * 8 * K)
    achieved = alg_bytes / (ms / 1000.0) / 1e9
    log(f"headline ({headline['mode']}, ef {headline['ef']}, {rows} rows): recall {rec:.4f}, {ms:.3f} ms, "
        f"{qps:,.0f} QPS; {expansions / nq:.1f} expansions, {(distances - dups) / nq:.1f} distances, "
        f"{alg_bytes / nq / 1e3:.1f} KB per query -> {achieved:.0f} GB/s")
    f32_leg = None
    if rows == "u8":
        # the same launches over the FP32 rows: the kernel the row copy replaces
        f32 = nat.SearchOpts()
        f32.row_format = 1
        ms32, _ = run.timed(headline["ef"], args.steps, args.warmup, f32)
        res32 = run.results()
        counts32 = run.counters(headline["ef"], f32, visited_log2=13)
        bytes32 = (nq * 4 * d + counts32[2] * 4 * (1 + M) + counts32[0] * 4 * (1 + 2 * M)
                   + (counts32[1] - counts32[3]) * 4 * d + nq * 8 * K)
        f32_leg = {"ms": round(ms32, 4), "qps": round(nq / ms32 * 1000, 1),
                   "identical_to_u8_rows": bool(res32.tobytes() == res.tobytes()),
                   "achieved_gbs": round(bytes32 / (ms32 / 1000.0) / 1e9, 1),
                   "frac": round(bytes32 / (ms32 / 1000.0) / 1e9 / HBM_PEAK_GBS, 4),
                   "traffic": measured_traffic(headline["mode"], headline["ef"], nq, "f32"),
                   "algorithmic_bytes_per_launch": bytes32}
        log(f"same ef over the FP32 rows: {ms32:.3f} ms, {nq / ms32 * 1000:,.0f} QPS "
            f"({ms32 / ms:.2f}x the 8-bit rows' time), results identical: {f32_leg['identical_to_u8_rows']}")

    # The host entry point at the same ef (HnswIndex::search_batch's shape: queries from and
    # results to host memory, so PCIe is inside the time): warmup + median of 5.
    index.search_arrays(queries, K, headline["ef"], opts=headline["opts"])
    host_times = []
    for _ in range(5):
        t = time.perf_counter()
        index.search_arrays(queries, K, headline["ef"], opts=headline["opts"])
        host_times.append(time.perf_counter() - t)
    host_ms = sorted(host_times)[2] * 1000
    log(f"host entry point (PCIe inclusive), ef {headline['ef']}: {host_ms:.3f} ms, "
        f"{nq / host_ms * 1000:,.0f} QPS")

    cpu = None
    identical = None
    if not args.no_cpu:
        from oracle import oracle as o
        g = index.graph_arrays()
        ef = fixed_best["ef"]
        o.hnsw_search(g, queries, K, ef, threads=threads)
        times = []
        for _ in range(5):
            t = time.perf_counter()
            ref, _ = o.hnsw_search(g, queries, K, ef, threads=threads)
            times.append(time.perf_counter() - t)
        med = sorted(times)[len(times) // 2]
        run.launch(ef)
        gpu_fixed = run.results()
        identical = float(np.mean([gpu_fixed[q].tobytes() == ref[q].tobytes() for q in range(nq)]))
        cpu = {"value": round(nq / med, 1), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"all {nq} queries at fixed ef={ef} (recall {fixed_best['recall']}), warmup + "
                         f"median of 5 passes (vecdex bench/hnsw_eval.cpp:493-512), "
                         f"{os.cpu_count()} host cpus, {threads} threads"}
        log(f"CPU baseline: {nq / med:,.0f} QPS on {threads} threads; GPU fixed-ef results identical "
            f"to the CPU's for {identical * 100:.2f}% of queries")

    line = {
        "metric": "QPS @ recall@10=0.95, 1M×128 L2, 1 GPU: learned vs fixed-ef vs host CPU",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"GPU HNSW search ({headline['mode']}): {args.n:,} x 128 L2 Gaussian "
                               f"mixture (1,024 components, integer-valued), {nq:,} queries, graph "
                               f"M=16 efConstruction=200 seed 0, k=10",
                   "ef": headline["ef"], "termination": headline["mode"],
                   "recall@10": round(rec, 4), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(headline["mode"], headline["ef"], nq, rows),
                     "algorithmic_bytes_per_launch": alg_bytes, "rows": rows},
        "cpu_baseline": cpu,
        "fixed_ef": {"ef": fixed["ef"], "recall": fixed["recall"], "qps": fixed["qps"],
                     "best_fine_ef": fixed_best["ef"], "best_fine_qps": fixed_best["qps"],
                     "best_fine_recall": fixed_best["recall"]},
        "headline_over_f32_rows": f32_leg,
        "learned": learned_leg,
        "sweep": sweep,
        "bruteforce": bf_leg,
        "serving": serving_leg,
        "gpu_build": gpu_build_leg,
        "load": load_leg,
        "base_10m": base10m_leg,
        "host_entry_point": {"ms": round(host_ms, 4), "qps": round(nq / host_ms * 1000, 1),
                             "note": "queries from and results to host memory (PCIe inclusive)"},
        "gpu_cpu_identical_fraction": identical,
        "wall_ms_per_step": round(wall_ms, 4),
        "ground_truth_s": round(gt_s, 3),
        "graph_build_s": None if build_s is None else round(build_s, 1),
    }
    print(json.dumps(line), flush=True)
    return 0


_COS = {}


def cosine_setup(args, vdx, datasets, threads):
    """The 1.2M x 100 cosine base (anisotropic Gaussian) and its graph, made once."""
    if not _COS:
        gen = datasets.AnisotropicGaussian(100, 0.8, seed=0)
        base = gen.sample(args.n_cos, 1)
        index, _ = build_or_load_graph(vdx, "aniso1200k_cos", base, vdx.Metric.Cosine, threads)
        _COS.update(gen=gen, base=base, index=index)
    return _COS["gen"], _COS["base"], _COS["index"]


def run_serving(args, vdx, nat, datasets, threads):
    """Uneven batches on the cosine base: 65,536 queries with ef ~ U{16, 32, 64, 128, 256, 512}
    and k ~ U{1, 10, 100}. One launch sized for the largest (ef, k) against the class-batched
    path (eager, then tuned and replayed as a HIP graph), results compared byte for byte; then
    batch sizes 1 .. 65,536 through the tuned graph."""
    gen, base, index = cosine_setup(args, vdx, datasets, threads)
    bf = vdx.BruteForceIndex(base.shape[1], vdx.Metric.Cosine)
    bf.add(base)
    bf_cos = bruteforce_leg(args, nat, bf, prepared(nat, gen.sample(10_000, 9), True), vdx.Metric.Cosine, threads)
    del bf
    nq = args.serving_nq
    rng = np.random.default_rng(5)
    efs = np.array([16, 32, 64, 128, 256, 512], dtype=np.uint32)[rng.integers(0, 6, nq)]
    ks = np.array([1, 10, 100], dtype=np.uint32)[rng.integers(0, 3, nq)]
    queries = prepared(nat, gen.sample(nq, 5), True)
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(queries, ctx)
    dks = nat.DeviceArray.from_numpy(ks, ctx)
    defs = nat.DeviceArray.from_numpy(efs, ctx)
    out = nat.DeviceArray((nq, 100), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(nq, np.uint32, ctx)
    handle = index.gpu_handle()

    def timed(fn, reps):
        fn()
        nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
        nat.check(nat.lib.vdx_gpu_event_record(ctx.handle, 2))
        for _ in range(reps):
            fn()
        nat.check(nat.lib.vdx_gpu_event_record(ctx.handle, 3))
        nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
        return nat.elapsed_ms(ctx, 2, 3) / reps

    def single(n=nq):
        nat.check(nat.lib.vdx_gpu_hnsw_search_device(handle, dq.ptr, n, dks.ptr, 0, defs.ptr, 512, 100,
                                                     out.ptr, 100, cnt.ptr, None, None))

    reps = max(3, args.steps // 4)
    ms_single = timed(single, reps)
    ref = out.numpy()
    batch = vdx.SearchBatch(index, nq, 100, 512)
    ms_eager = timed(lambda: batch.run(dq, nq, out, cnt, ks=dks, efs=defs, graph=False), reps)
    t = time.perf_counter()
    classes = batch.tune(dq, nq, out, cnt, ks=dks, efs=defs, reps=2)
    tune_s = time.perf_counter() - t
    ms_graph = timed(lambda: batch.run(dq, nq, out, cnt, ks=dks, efs=defs, graph=True), reps)
    identical = bool(out.numpy().tobytes() == ref.tobytes())
    log(f"serving {nq} mixed queries: one launch {ms_single:.2f} ms ({nq / ms_single * 1000:,.0f} QPS), "
        f"classes eager {ms_eager:.2f} ms, tuned graph {ms_graph:.2f} ms "
        f"({nq / ms_graph * 1000:,.0f} QPS); identical results: {identical}; tuning {tune_s:.1f} s")
    for c in classes:
        log(f"  class cap {c['ef_cap']}: {c['queries']} queries, visited 2^{c['visited_log2']}, "
            f"{c['waves_per_group']} waves/group, blocks/CU {c['blocks_per_cu'] or 'max'}, {c['ms']:.2f} ms")
    sizes = []
    for b in sorted({*(2 ** i for i in range(17) if 2 ** i < nq), nq}):
        r = 20 if b <= 1024 else 5
        ms = timed(lambda: batch.run(dq, b, out, cnt, ks=dks, efs=defs, graph=True), r)
        sizes.append({"batch": b, "ms": round(ms, 4), "qps": round(b / ms * 1000, 1)})
        log(f"  batch {b}: {ms:.3f} ms per call, {b / ms * 1000:,.0f} QPS")
    del batch
    return {"workload": f"{nq} queries on the {args.n_cos:,} x 100 cosine base, ef ~ U{{16..512}}, "
                        f"k ~ U{{1, 10, 100}}", "one_launch_ms": round(ms_single, 3),
            "classes_eager_ms": round(ms_eager, 3), "tuned_graph_ms": round(ms_graph, 3),
            "tuned_graph_qps": round(nq / ms_graph * 1000, 1),
            "speedup_vs_one_launch": round(ms_single / ms_graph, 3),
            "identical_to_one_launch": identical, "tune_s": round(tune_s, 2),
            "classes": classes, "batch_sizes": sizes, "bruteforce_cosine_1200k_x_100": bf_cos}


def run_10m(args, vdx, nat, datasets, threads):
    """The 10,000,000 x 128 L2 base (the 1M generator with 4,096 components, graph M=32,
    efConstruction=200): fixed-ef recall and QPS over 10,000 queries, then 65,536 queries with
    ef ~ U{16..512} and k ~ U{1, 10, 100} through the tuned class graph against one launch."""
    t = time.perf_counter()
    mix = datasets.GaussianMixture(128, 4096, 0.0, 128.0, 4.0, 32.0, mixture_seed=0)
    base = mix.sample(args.n10m, 1)
    queries = mix.sample(10_000, 2)
    log(f"10M: generated {len(base):,} x 128 in {time.perf_counter() - t:.1f} s")
    CACHE.mkdir(parents=True, exist_ok=True)
    path = CACHE / f"gmm4096_n{len(base)}_M32_efc200_seed0.vdx"
    build_s = None
    if path.exists():
        index = vdx.HnswIndex.load(path)
    else:
        index = vdx.HnswIndex(128, vdx.Metric.L2, M=32, ef_construction=200, seed=0)
        t = time.perf_counter()
        step = 1_000_000
        how = "host" if args.host_build_10m else "GPU-assisted"
        for r0 in range(0, len(base), step):  # in slices, reporting progress
            index.add(base[r0:r0 + step], num_threads=threads, gpu_assisted=not args.host_build_10m)
            split = ""
            if not args.host_build_10m:
                st = index.last_build_stats
                split = (f"; this slice: host {st['host_insert_s']:.1f} s, sync {st['sync_s']:.1f} s, "
                         f"search {st['search_s']:.1f} s, link {st['link_s']:.1f} s")
            log(f"10M: {min(r0 + step, len(base)):,} rows linked ({how}), {time.perf_counter() - t:.0f} s{split}")
        build_s = time.perf_counter() - t
        log(f"10M: built graph (M=32, efC=200, {how}) on {threads} threads in {build_s:.0f} s")
        if "VDX_BENCH_CACHE" in os.environ:  # an 8 GB file: kept only where asked for
            index.save(path)
    t = time.perf_counter()
    truth, bf = ground_truth(vdx, base, vdx.Metric.L2, queries)
    log(f"10M: ground truth (BF16 brute force, certified) in {time.perf_counter() - t:.1f} s")
    del bf
    run = Runner(nat, index, queries)
    sweep = []
    for ef in (16, 32, 48, 64, 96, 128, 256):
        run.launch(ef)
        r = recall(run.results(), truth)
        ms, _ = run.timed(ef, args.steps, args.warmup)
        sweep.append({"ef": ef, "recall": round(r, 4), "ms": round(ms, 4), "qps": round(10_000 / ms * 1000, 1)})
        log(f"10M fixed ef {ef}: recall {r:.4f}, {ms:.3f} ms, {sweep[-1]['qps']:,.0f} QPS")
    at = [x for x in sweep if x["recall"] >= TARGET_RECALL]
    nq = args.serving_nq
    rng = np.random.default_rng(6)
    efs = np.array([16, 32, 64, 128, 256, 512], dtype=np.uint32)[rng.integers(0, 6, nq)]
    ks = np.array([1, 10, 100], dtype=np.uint32)[rng.integers(0, 3, nq)]
    sq = mix.sample(nq, 7)
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(sq, ctx)
    dks = nat.DeviceArray.from_numpy(ks, ctx)
    defs = nat.DeviceArray.from_numpy(efs, ctx)
    out = nat.DeviceArray((nq, 100), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(nq, np.uint32, ctx)
    handle = index.gpu_handle()

    def timed(fn, reps):
        fn()
        nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
        nat.check(nat.lib.vdx_gpu_event_record(ctx.handle, 2))
        for _ in range(reps):
            fn()
        nat.check(nat.lib.vdx_gpu_event_record(ctx.handle, 3))
        nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
        return nat.elapsed_ms(ctx, 2, 3) / reps

    ms_one = timed(lambda: nat.check(nat.lib.vdx_gpu_hnsw_search_device(
        handle, dq.ptr, nq, dks.ptr, 0, defs.ptr, 512, 100, out.ptr, 100, cnt.ptr, None, None)), 3)
    ref = out.numpy()
    batch = vdx.SearchBatch(index, nq, 100, 512)
    batch.tune(dq, nq, out, cnt, ks=dks, efs=defs, reps=1)
    ms_graph = timed(lambda: batch.run(dq, nq, out, cnt, ks=dks, efs=defs, graph=True), 3)
    identical = bool(out.numpy().tobytes() == ref.tobytes())
    log(f"10M serving {nq} mixed queries: one launch {ms_one:.1f} ms, tuned graph {ms_graph:.1f} ms "
        f"({nq / ms_graph * 1000:,.0f} QPS), identical {identical}")
    del batch
    return {"workload": f"{len(base):,} x 128 L2 Gaussian mixture (4,096 components), M=32, efC=200",
            "graph_build_s": None if build_s is None else round(build_s, 1),
            "graph_build": "host" if args.host_build_10m else "GPU-assisted", "fixed_ef_sweep": sweep,
            "qps_at_recall_0.95": at[0] if at else None,
            "serving": {"queries": nq, "one_launch_ms": round(ms_one, 2), "tuned_graph_ms": round(ms_graph, 2),
                        "tuned_graph_qps": round(nq / ms_graph * 1000, 1), "identical_to_one_launch": identical}}


def bruteforce_leg(args, nat, bf, queries, metric, threads):
    """BruteForceIndex::search on the GPU over all (prepared) queries at k=10: the exact scan (the
    reference's result), then the FP32 and BF16 MFMA shortlists with exact re-rank, each checked
    byte for byte against the exact scan; plus the host's exact scan (the oracle's restatement of
    the reference, on all host threads) over a bounded sample of the queries."""
    base = bf.vectors()  # as stored: normalized under cosine
    nq, (n, d) = len(queries), base.shape
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(np.ascontiguousarray(queries, dtype=np.float32), ctx)
    bout = nat.DeviceArray((nq, K), nat.NEIGHBOR_DTYPE, ctx)
    bcnt = nat.DeviceArray(nq, np.uint32, ctx)
    leg, exact = {}, None
    flops = 2.0 * nq * n * d
    for prec in ("exact", "fp32", "bf16"):
        st = nat.BfStats()
        times = []
        for _ in range(1 if prec == "exact" else 3):
            nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
            t0 = time.perf_counter()
            nat.check(nat.lib.vdx_gpu_bf_search_device(bf.gpu_handle(), dq.ptr, nq, K, nat.PRECISIONS[prec],
                                                       bout.ptr, K, bcnt.ptr, ctypes.byref(st)))
            times.append((time.perf_counter() - t0) * 1000)
        res = bout.numpy()
        ms = min(times)
        if prec == "exact":
            exact = res
            leg["exact"] = {"ms": round(ms, 2), "qps": round(nq / ms * 1000, 1)}
            log(f"brute force exact scan ({n:,} x {d}): {ms:.1f} ms per {nq:,} queries")
            continue
        peak = FP32_MFMA_TFLOPS if prec == "fp32" else BF16_MFMA_TFLOPS
        leg[prec] = {"ms": round(ms, 2), "qps": round(nq / ms * 1000, 1),
                     "identical_to_exact": bool(res.tobytes() == exact.tobytes()),
                     "exact_fallbacks": int(st.exact_fallback), "shortlist": int(st.shortlist),
                     "end_to_end_tflops": round(flops / ms / 1e9, 1),
                     "frac_of_mfma_peak": round(flops / ms / 1e9 / peak, 4)}
        log(f"brute force {prec}: {ms:.1f} ms per {nq:,} queries, identical to the exact scan "
            f"{leg[prec]['identical_to_exact']}, fallbacks {st.exact_fallback}")
    leg["bf16_speedup_vs_fp32"] = round(leg["fp32"]["ms"] / leg["bf16"]["ms"], 2)
    if not args.no_cpu:
        from oracle import oracle as o
        sample = 200
        t0 = time.perf_counter()
        host = o.bruteforce_f32(base, np.ascontiguousarray(queries[:sample]), K, metric=metric, threads=threads)
        secs = time.perf_counter() - t0
        leg["cpu_exact"] = {"qps": round(sample / secs, 1), "cores": threads, "kind": "port",
                            "sample": f"{sample} of the {nq:,} queries, oracle exact scan (FP32, AVX2 order)",
                            "identical_ids": bool(np.array_equal(host["id"], exact["id"][:sample]))}
        log(f"host exact scan: {sample / secs:,.0f} QPS on {threads} threads (sample of {sample}); GPU/host "
            f"speed-up of the BF16 path {leg['bf16']['qps'] / (sample / secs):,.0f}x")
    del dq, bout, bcnt
    return leg


def run_load(vdx, index, queries):
    """Index file to a searchable HBM copy, 1M graph: FORMAT v1 through the host graph (load,
    then upload) against FORMAT v2 streamed straight to HBM. The files were just written, so the
    page cache is warm: this times parsing, checks and copies, not the disk."""
    CACHE.mkdir(parents=True, exist_ok=True)
    p1, p2 = CACHE / "load_test.vdx", CACHE / "load_test.vdx2"
    index.save(p1)
    index.save(p2, version=2)
    out = {"bytes_v1": p1.stat().st_size, "bytes_v2": p2.stat().st_size}
    ref, _ = index.search_arrays(queries[:1000], K, 64)
    for name, fn in (("v1_host_then_upload", lambda: vdx.HnswIndex.load(p1)),
                     ("v2_host_then_upload", lambda: vdx.HnswIndex.load(p2)),
                     ("v2_device", lambda: vdx.HnswDeviceIndex.load(p2))):
        t = time.perf_counter()
        loaded = fn()
        loaded.gpu_handle()
        secs = time.perf_counter() - t
        got, _ = loaded.search_arrays(queries[:1000], K, 64)
        out[name] = {"s": round(secs, 3), "identical_results": bool(got.tobytes() == ref.tobytes())}
        log(f"load {name}: {secs:.2f} s, results identical {out[name]['identical_results']}")
        del loaded
    p1.unlink()
    p2.unlink()
    return out


def run_gpu_build(vdx, nat, base, queries, truth, threads, host_build_s, sweep):
    """The same 1M graph (M=16, efC=200, seed 0) built GPU-assisted: level-0 insertion searches
    in batches on the GPU, selection and linking on the host threads; build time, its split, and
    recall/QPS of the resulting graph against the host-built one's at the same ef."""
    t = time.perf_counter()
    g = vdx.HnswIndex(base.shape[1], vdx.Metric.L2, M=16, ef_construction=200, seed=0)
    g.add(base, num_threads=threads, gpu_assisted=True)
    secs = time.perf_counter() - t
    st = g.last_build_stats
    log(f"GPU-assisted build: {secs:.1f} s on {threads} threads (host build {host_build_s} s); "
        + ", ".join(f"{k} {v:.2f}" if isinstance(v, float) else f"{k} {v}" for k, v in st.items()))
    run = Runner(nat, g, queries)
    rows = []
    host_rows = {r["ef"]: r for r in sweep}
    for ef in (36, 64):
        run.launch(ef)
        r = recall(run.results(), truth)
        ms, _ = run.timed(ef, 10, 2)
        host = host_rows.get(ef)
        rows.append({"ef": ef, "recall": round(r, 4), "qps": round(len(queries) / ms * 1000, 1),
                     "host_graph_recall": host["recall"] if host else None})
        log(f"GPU-built graph, ef {ef}: recall {r:.4f} (host-built {rows[-1]['host_graph_recall']}), "
            f"{rows[-1]['qps']:,.0f} QPS")
    return {"build_s": round(secs, 2), "host_build_s": None if host_build_s is None else round(host_build_s, 2),
            "threads": threads, "stats": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()},
            "search": rows}


def run_learned(args, vdx, nat, datasets, learned, index, base, queries, truth, mix, threads,
                fixed, fixed_best, headline):
    """Traces, training, calibration and evaluation of the learned stop on both bases.

    Traces: half the training queries of each base at each of its two trace caps (L2: ef 64
    and 128; cosine: 128 and 256), labelled by brute-force top-10 hits. One model for both.
    Calibration (held-out queries, never the evaluation set): the recall target is 0.95 on the
    L2 base and, on the cosine base (where 0.95 needs ef far above 512), the recall fixed ef=128
    reaches; for each ef cap the threshold with the fewest expansions reaching target + 0.005,
    then the cap with the highest calibration QPS. The evaluation set then measures recall and
    QPS against the fixed-ef operating point."""
    t_all = time.perf_counter()
    cos_gen, cos_base, cos_index = cosine_setup(args, vdx, datasets, threads)
    ntr = args.train_queries
    sets = {
        "l2": {"index": index, "base": base, "metric": vdx.Metric.L2, "trace_efs": (64, 128),
               "caps": (48, 64, 96, 128),
               "train": mix.sample(ntr, 3), "calib": mix.sample(args.nq, 4), "eval": queries},
        "cos": {"index": cos_index, "base": cos_base, "metric": vdx.Metric.Cosine,
                "trace_efs": (128, 256), "caps": (128, 192, 256),
                "train": prepared(nat, cos_gen.sample(ntr, 3), True),
                "calib": prepared(nat, cos_gen.sample(args.nq, 4), True),
                "eval": prepared(nat, cos_gen.sample(args.nq, 2), True)},
    }
    feats, labels = [], []
    for name, s in sets.items():
        allq = np.concatenate([s["train"], s["calib"], s["eval"]])
        t = time.perf_counter()
        tr, _ = ground_truth(vdx, s["base"], s["metric"], allq)
        s["truth_train"] = tr[:ntr]
        s["truth_calib"] = tr[ntr:ntr + args.nq]
        s["truth_eval"] = truth if name == "l2" else tr[ntr + args.nq:]
        log(f"learned/{name}: ground truth for {len(allq)} queries in {time.perf_counter() - t:.1f} s")
        half = ntr // 2
        for part, ef in enumerate(s["trace_efs"]):
            lo, hi = part * half, (part + 1) * half
            t = time.perf_counter()
            f, lab, steps, final = learned.collect_traces(s["index"], s["train"][lo:hi],
                                                          s["truth_train"]["id"][lo:hi], k=K, ef=ef,
                                                          max_steps=2 * ef)
            log(f"learned/{name}: traces at ef {ef}: {len(f):,} steps ({steps.mean():.1f} "
                f"expansions/query, label mean {lab.mean():.3f}) in {time.perf_counter() - t:.1f} s")
            feats.append(f)
            labels.append(lab)
    f = np.concatenate(feats)
    lab = np.concatenate(labels)
    if len(f) > args.max_train_steps:
        pick = np.random.default_rng(0).choice(len(f), args.max_train_steps, replace=False)
        f, lab = f[pick], lab[pick]
    t = time.perf_counter()
    model, losses = learned.train_stop_model(f, lab, epochs=20, batch=4096, lr=1e-3, seed=0)
    train_s = time.perf_counter() - t
    log(f"learned: trained on {len(f):,} steps, 20 epochs in {train_s:.1f} s; loss "
        f"{losses[0]:.4f} -> {losses[-1]:.4f}")

    results = {}
    for name, s in sets.items():
        rc = Runner(nat, s["index"], s["calib"])
        re_ = Runner(nat, s["index"], s["eval"])
        if name == "l2":
            target = TARGET_RECALL
            fixed_eval = {"ef": fixed_best["ef"], "recall": fixed_best["recall"], "qps": fixed_best["qps"]}
        else:
            rc.launch(128)
            target = recall(rc.results(), s["truth_calib"])
            fixed_eval = None
            sweep_cos = []
            for ef in (64, 96, 128, 192, 256):
                re_.launch(ef)
                r_e = recall(re_.results(), s["truth_eval"])
                ms, _ = re_.timed(ef, args.steps, args.warmup)
                sweep_cos.append({"ef": ef, "recall": round(r_e, 4), "qps": round(len(s["eval"]) / ms * 1000, 1)})
                log(f"cos fixed ef {ef}: recall {r_e:.4f}, {sweep_cos[-1]['qps']:,.0f} QPS")
                if fixed_eval is None and r_e >= target:
                    fixed_eval = sweep_cos[-1]
            results["cos_fixed_sweep"] = sweep_cos
        best = None
        for ef_cap in s["caps"]:
            thr, rec_c, work, table = learned.calibrate(s["index"], model, s["calib"],
                                                        s["truth_calib"]["id"], K, ef_cap, target + 0.005)
            if rec_c < target + 0.005:
                log(f"learned/{name} ef cap {ef_cap}: no threshold reaches {target + 0.005:.4f} "
                    f"(best {rec_c:.4f})")
                continue
            opts = learned.learned_opts(model, thr, 1)
            ms_c, _ = rc.timed(ef_cap, args.steps, args.warmup, opts)
            log(f"learned/{name} ef cap {ef_cap}: threshold {thr}, calib recall {rec_c:.4f}, "
                f"{work:.1f} expansions, {len(s['calib']) / ms_c * 1000:,.0f} QPS on calibration")
            if best is None or ms_c < best["_ms_c"]:
                best = {"ef_cap": ef_cap, "threshold": thr, "calib_recall": round(rec_c, 4),
                        "_ms_c": ms_c, "_opts": opts}
        if best is not None:
            opts = best["_opts"]
            re_.launch(best["ef_cap"], opts)
            rec_e = recall(re_.results(), s["truth_eval"])
            ms, _ = re_.timed(best["ef_cap"], args.steps, args.warmup, opts)
            exp = re_.counters(best["ef_cap"], opts)[0] / len(s["eval"])
            best.update({"target": round(target, 4), "eval_recall": round(rec_e, 4), "ms": round(ms, 4),
                         "qps": round(len(s["eval"]) / ms * 1000, 1), "expansions": round(exp, 1),
                         "fixed_ef_at_target": fixed_eval})
            if fixed_eval is not None:
                best["speedup_vs_fixed_ef"] = round(best["qps"] / fixed_eval["qps"], 3)
            log(f"learned/{name}: eval recall {rec_e:.4f} (target {target:.4f}), {ms:.3f} ms, "
                f"{best['qps']:,.0f} QPS, {exp:.1f} expansions; fixed ef at target: {fixed_eval}")
        results[name] = best
    l2 = results["l2"]
    leg = {
        "train_steps": len(f), "train_s": round(train_s, 1),
        "loss_first_last": [round(float(losses[0]), 4), round(float(losses[-1]), 4)],
        "l2": {k: v for k, v in (l2 or {}).items() if not k.startswith("_")},
        "cos": {k: v for k, v in (results["cos"] or {}).items() if not k.startswith("_")},
        "cos_fixed_sweep": results["cos_fixed_sweep"],
        "total_s": round(time.perf_counter() - t_all, 1),
    }
    if l2 is not None and l2["eval_recall"] >= TARGET_RECALL and l2["qps"] > headline["qps"]:
        headline = {"mode": "learned", "ef": l2["ef_cap"], "opts": l2["_opts"], "qps": l2["qps"],
                    "recall": l2["eval_recall"], "model": model}  # the model must stay alive
    return leg, headline


if __name__ == "__main__":
    sys.exit(main())
