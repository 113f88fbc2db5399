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
FP8_MFMA_TFLOPS = 5000.0
EF_SWEEP = (16, 32, 64, 128, 512)
EF_FINE = (34, 36, 38, 40, 44, 48, 56)
BUDGET_EFS = (36, 38, 40, 42, 44, 48, 56)    # the expansion-budget leg's grid
BUDGETS = (44, 48, 52, 56, 60, 64, 72, 96)
TARGET_RECALL = 0.95
CALIB_MARGIN = 0.002  # about the sampling error of recall@10 over 10,000 held-out queries
THRESHOLDS = (0.5, 0.6, 0.65, 0.7, 0.72, 0.75, 0.78, 0.8, 0.83, 0.85, 0.88, 0.9, 0.93, 0.95, 0.97, 0.98, 0.99,
              0.995, 0.999)
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


def timed_best_form(run, ef, steps, warmup, opts=None):
    """run.timed with the faster visited-set form where there is a choice (R capacity above 512,
    the LDS beam: the LDS cache or the HBM bitmap; the same results either way). Returns
    (ms, form), form 0 where the default is the only one."""
    if ef <= 512:
        ms, _ = run.timed(ef, steps, warmup, opts)
        return ms, 0
    best = None
    for form in (1, 2):
        o = run.nat.SearchOpts.from_buffer_copy(opts) if opts is not None else run.nat.SearchOpts()
        o.visited_form = form
        ms, _ = run.timed(ef, steps, warmup, o)
        if best is None or ms < best[0]:
            best = (ms, form)
    return best


def recall(found, truth):
    """Mean recall@10 by id (vecdex src/recall.cpp:11-34, through the library)."""
    from hnsw_vector_search_amd import recall as r
    return r.mean_recall_at_k(found, truth["id"].astype(np.int32), K)


def threshold_recall(found, truth):
    """Mean distance-threshold recall@10 (src/recall.cpp:36-51, through the library)."""
    from hnsw_vector_search_amd import recall as r
    return r.mean_threshold_recall_at_k(found, truth["distance"][:, K - 1], K)


def measured_traffic(mode, ef, nq, rows, max_expansions=0):
    """HBM bytes per launch of the headline kernel from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), when it covers this
    workload (rows: "u8" for the 8-bit row copy, "f32"; the expansion budget, 0 for none);
    else None."""
    path = ROOT / "profiles" / "hnsw_beam_traffic.json"
    if not path.exists():
        return None
    for row in json.loads(path.read_text()).get("workloads", []):
        if (row.get("mode") == mode and row.get("ef") == ef and row.get("nq") == nq
                and row.get("rows", "f32") == rows and row.get("max_expansions", 0) == max_expansions):
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
    ap.add_argument("--no-base10m", action="store_true",
                    help="skip the 10M x 128 M=32 base (its GPU-assisted graph build takes about half a minute)")
    ap.add_argument("--n10m", type=int, default=10_000_000)
    ap.add_argument("--no-gpu-build", action="store_true", help="skip the GPU-assisted build leg")
    ap.add_argument("--no-load", action="store_true", help="skip the FORMAT v1/v2 load leg")
    ap.add_argument("--host-build-10m", action="store_true",
                    help="build the 10M graph on the host alone (default: GPU-assisted)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-bruteforce", action="store_true", help="skip the brute-force leg (profiling runs)")
    ap.add_argument("--efs", type=str, default=",".join(map(str, EF_SWEEP + EF_FINE)))
    ap.add_argument("--budget-efs", type=str, default=",".join(map(str, BUDGET_EFS)),
                    help="ef grid of the expansion-budget leg")
    ap.add_argument("--budgets", type=str, default=",".join(map(str, BUDGETS)),
                    help="expansion budgets of the budget leg (one ef and one budget pin the policy)")
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

    # Ground truth (and the brute-force leg): exact scan, FP32 vs BF16 MFMA shortlist. The
    # held-out calibration queries (same mixture, their own seed) choose the fixed ef; the
    # evaluation queries then measure it.
    t = time.perf_counter()
    truth, bf = ground_truth(vdx, base, vdx.Metric.L2, queries)
    gt_s = time.perf_counter() - t
    calib_queries = mix.sample(args.nq, 4)
    calib_truth, _ = bf.search_arrays(calib_queries, K, "bf16")
    bf_leg = None if args.no_bruteforce else bruteforce_leg(args, nat, bf, queries, vdx.Metric.L2, threads)
    del bf

    # Fixed-ef sweep: recall on both query sets, time on the evaluation set.
    run = Runner(nat, index, queries)
    run_calib = Runner(nat, index, calib_queries)
    sweep = []
    for ef in sorted(set(map(int, args.efs.split(",")))):
        run.launch(ef)
        res = run.results()
        run_calib.launch(ef)
        ms, _ = run.timed(ef, args.steps, args.warmup)
        row = {"ef": ef, "recall": round(recall(res, truth), 4),
               "recall_threshold": round(threshold_recall(res, truth), 4),
               "calib_recall": round(recall(run_calib.results(), calib_truth), 4),
               "ms": round(ms, 4), "qps": round(args.nq / ms * 1000, 1)}
        sweep.append(row)
        log(f"fixed ef {ef}: recall {row['recall']:.4f} (held-out {row['calib_recall']:.4f}), {ms:.3f} ms/pass, "
            f"{row['qps']:,.0f} QPS")
    del run_calib
    gpu_build_leg = None if args.no_gpu_build else run_gpu_build(vdx, nat, base, queries, truth, threads,
                                                                  build_s, sweep)
    load_leg = None if args.no_load else run_load(vdx, index, queries)
    # the operating points: the smallest ef whose recall on the held-out queries reaches 0.95
    config_points = [s for s in sweep if s["ef"] in EF_SWEEP and s["calib_recall"] >= TARGET_RECALL]
    fixed = min(config_points, key=lambda s: s["ef"]) if config_points else sweep[-1]
    fine_points = [s for s in sweep if s["calib_recall"] >= TARGET_RECALL]
    fixed_best = min(fine_points, key=lambda s: s["ef"]) if fine_points else fixed

    # Adaptive termination: the expansion budget, then the learned stop (with and without it).
    learned_leg = None
    headline = {"mode": "fixed-ef", "ef": fixed_best["ef"], "opts": None, "recall": fixed_best["recall"],
                "qps": fixed_best["qps"]}
    budget_leg, headline = run_budget(args, nat, index, queries, truth, calib_queries, calib_truth,
                                      fixed_best, headline)
    if not args.no_learned:
        learned_leg, headline = run_learned(args, vdx, nat, datasets, learned, index, base, queries,
                                            truth, mix, threads, fixed, fixed_best, headline, budget_leg)
    serving_leg = None if args.no_serving else run_serving(args, vdx, nat, datasets, threads)
    base10m_leg = None if args.no_base10m else run_10m(args, vdx, nat, datasets, threads)

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
    # The plain and learned searches read the 8-bit row copy when the index keeps one (every
    # stored value an integer in [0, 255]: the same floats, so the same results): d bytes per
    # row, else 4d.
    rows = "u8" if index.gpu_u8_rows() else "f32"
    row_bytes = d if rows == "u8" else 4 * d
    alg_bytes = (nq * 4 * d + upper_blocks * 4 * (1 + M) + expansions * 4 * (1 + 2 * M)
                 + (distances - dups) * row_bytes + nq * 8 * K)
    achieved = alg_bytes / (ms / 1000.0) / 1e9
    log(f"headline ({headline['mode']}, ef {headline['ef']}, {rows} rows): recall {rec:.4f}, {ms:.3f} ms, "
        f"{qps:,.0f} QPS; {expansions / nq:.1f} expansions, {(distances - dups) / nq:.1f} distances, "
        f"{alg_bytes / nq / 1e3:.1f} KB per query -> {achieved:.0f} GB/s")
    gather = gather_ceiling(nat, run.ctx, len(base), achieved)
    f32_leg = None
    if rows == "u8":
        # the same launches over the FP32 rows: the kernel the row copy replaces
        # (the headline's own options, the row format forced)
        f32 = nat.SearchOpts.from_buffer_copy(headline["opts"]) if headline["opts"] is not None else nat.SearchOpts()
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
                   "traffic": measured_traffic(headline["mode"], headline["ef"], nq, "f32", f32.max_expansions),
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

    single = run_single_query(args, nat, index, queries, threads)

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
                   "max_expansions": headline["opts"].max_expansions if headline["opts"] is not None else 0,
                   "recall@10": round(rec, 4), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(headline["mode"], headline["ef"], nq, rows,
                                                 headline["opts"].max_expansions if headline["opts"] else 0),
                     "algorithmic_bytes_per_launch": alg_bytes, "rows": rows},
        "cpu_baseline": cpu,
        "gather_ceiling": gather,
        "fixed_ef": {"ef": fixed["ef"], "recall": fixed["recall"], "qps": fixed["qps"],
                     "best_fine_ef": fixed_best["ef"], "best_fine_qps": fixed_best["qps"],
                     "best_fine_recall": fixed_best["recall"]},
        "headline_over_f32_rows": f32_leg,
        "budget": budget_leg,
        "learned": learned_leg,
        "sweep": sweep,
        "bruteforce": bf_leg,
        "serving": serving_leg,
        "gpu_build": gpu_build_leg,
        "load": load_leg,
        "base_10m": base10m_leg,
        "host_entry_point": {"ms": round(host_ms, 4), "qps": round(nq / host_ms * 1000, 1),
                             "note": "queries from and results to host memory (PCIe inclusive)"},
        "single_query": single,
        "float_data": float_figures(f32_leg, learned_leg),
        "gpu_cpu_identical_fraction": identical,
        "wall_ms_per_step": round(wall_ms, 4),
        "ground_truth_s": round(gt_s, 3),
        "graph_build_s": None if build_s is None else round(build_s, 1),
    }
    print(json.dumps(line), flush=True)
    return 0


def gather_ceiling(nat, ctx, n, achieved_gbs, reps=3):
    """The rate random 128-byte row gathers reach on this GPU with no dependent loads between
    them (vdx_gpu_gather_probe), over tables the size of the base's 8-bit rows (n x 128 B), of
    those rows with the layer-0 link blocks (n x (128 + 132) B, what the search reads), and of
    1 GiB (past the 256 MB last-level cache); 8,192 waves, 32 rows in flight each as in the
    search's row loads, and a wave sweep on the second table. Best of `reps` event-timed runs.
    The search's algorithmic rate over the second figure is the fraction of the practical
    gather ceiling it reaches (its loads also wait on each other, which the probe's do not)."""
    big = 1 << 30
    table = nat.DeviceArray(big // 4, np.uint32, ctx)
    sink = nat.DeviceArray(8192, np.uint32, ctx)
    rows_read = ctypes.c_uint64(0)

    def rate(table_bytes, waves, per_wave=4096):
        best = None
        for _ in range(reps):
            nat.check(nat.lib.vdx_gpu_event_record(ctx.handle, 0))
            nat.check(nat.lib.vdx_gpu_gather_probe(ctx.handle, table.ptr, table_bytes, 128, per_wave, waves,
                                                   sink.ptr, ctypes.byref(rows_read)))
            nat.check(nat.lib.vdx_gpu_event_record(ctx.handle, 1))
            nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
            ms = nat.elapsed_ms(ctx, 0, 1)
            best = ms if best is None else min(best, ms)
        return round(rows_read.value * 128 / (best / 1000.0) / 1e9, 1)

    rows_only = min(big, n * 128)
    with_blocks = min(big, n * (128 + 132))
    out = {"row_bytes": 128, "waves": 8192,
           "rows_table_bytes": rows_only, "rows_gbs": rate(rows_only, 8192),
           "rows_and_blocks_table_bytes": with_blocks, "rows_and_blocks_gbs": rate(with_blocks, 8192),
           "table_1gib_gbs": rate(big, 8192),
           "waves_sweep_rows_and_blocks": {w: rate(with_blocks, w) for w in (1024, 2048, 4096, 7168)}}
    out["search_fraction_of_ceiling"] = round(achieved_gbs / out["rows_and_blocks_gbs"], 3)
    log(f"gather ceiling (random 128 B rows, 8,192 waves): {out['rows_gbs']:.0f} GB/s over {rows_only >> 20} MB, "
        f"{out['rows_and_blocks_gbs']:.0f} GB/s over {with_blocks >> 20} MB, {out['table_1gib_gbs']:.0f} GB/s over 1 GiB; "
        f"waves {out['waves_sweep_rows_and_blocks']}; the search reaches {out['search_fraction_of_ceiling']:.2f} of it")
    del table, sink
    return out


def run_single_query(args, nat, index, queries, threads, efs=(36, 64)):
    """HnswIndex::search of one query per call (src/hnsw.cpp:484-488): through the host entry
    point (validation, copies both ways, launch; mean over 300 calls), and one device-resident
    query per launch (HIP events over 200 launches); beside it the oracle's search of one query
    at a time on one host thread (vecdex's single-thread figure) over 2,000 queries."""
    out = {}
    for ef in efs:
        index.search_arrays(queries[:1], K, ef)
        t = time.perf_counter()
        for i in range(300):
            index.search_arrays(queries[i:i + 1], K, ef)
        host_us = (time.perf_counter() - t) / 300 * 1e6
        one = Runner(nat, index, queries[:1])
        ms1, _ = one.timed(ef, 200, 10)
        row = {"host_entry_us": round(host_us, 1), "device_launch_us": round(ms1 * 1000, 1)}
        if not args.no_cpu:
            from oracle import oracle as o
            g = index.graph_arrays()
            o.hnsw_search(g, queries[:200], K, ef, threads=1)
            t = time.perf_counter()
            o.hnsw_search(g, queries[:2000], K, ef, threads=1)
            row["cpu_one_thread_us"] = round((time.perf_counter() - t) / 2000 * 1e6, 1)
        out[f"ef{ef}"] = row
        log(f"single query, ef {ef}: host entry {host_us:.0f} us, device launch {ms1 * 1000:.0f} us, "
            f"CPU one thread {row.get('cpu_one_thread_us')} us")
        del one
    return out


def float_figures(f32_leg, learned_leg):
    """The figures that hold for any float data (no byte-row copy): the headline ef over the
    1M base's FP32 rows, and the 1.2M x 100 cosine base (float vectors) at recall 0.95."""
    out = {}
    if f32_leg is not None:
        out["fp32_rows_1M_x_128_same_ef"] = {"qps": f32_leg["qps"], "ms": f32_leg["ms"], "frac": f32_leg["frac"]}
    if learned_leg is not None:
        at = [r for r in learned_leg["cos_fixed_sweep"] if r["calib_recall"] >= TARGET_RECALL]
        if at:
            out["cosine_1_2M_x_100_at_0_95"] = at[0]
    return out


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
            f"{c['waves_per_group']} waves/group, blocks/CU {c['blocks_per_cu'] or 'max'}, "
            f"{'HBM bitmap' if c.get('visited_form') == 2 else 'LDS cache'}, {c['ms']:.2f} ms")
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
    # held-out queries of the same mixture choose the operating points, as on the 1M base
    calib = mix.sample(10_000, 4)
    t = time.perf_counter()
    both, bf = ground_truth(vdx, base, vdx.Metric.L2, np.concatenate([queries, calib]))
    truth, calib_truth = both[:len(queries)], both[len(queries):]
    log(f"10M: ground truth (BF16 brute force, certified) in {time.perf_counter() - t:.1f} s")
    del bf
    run = Runner(nat, index, queries)
    run_c = Runner(nat, index, calib)
    sweep = []
    for ef in (16, 32, 48, 64, 72, 80, 96, 128, 256):
        run.launch(ef)
        r = recall(run.results(), truth)
        run_c.launch(ef)
        r_c = recall(run_c.results(), calib_truth)
        ms, _ = run.timed(ef, args.steps, args.warmup)
        sweep.append({"ef": ef, "recall": round(r, 4), "calib_recall": round(r_c, 4), "ms": round(ms, 4),
                      "qps": round(10_000 / ms * 1000, 1)})
        log(f"10M fixed ef {ef}: recall {r:.4f} (held-out {r_c:.4f}), {ms:.3f} ms, {sweep[-1]['qps']:,.0f} QPS")
    at = [x for x in sweep if x["calib_recall"] >= TARGET_RECALL]
    # the expansion budget (as on the 1M base): (ef, budget) chosen on the held-out queries
    budget_leg = None
    best = None
    for budget in (72, 80, 96, 112, 128, 160):
        for ef in (64, 68, 72, 80):
            o = nat.SearchOpts()
            o.max_expansions = budget
            run_c.launch(ef, o)
            r_c = recall(run_c.results(), calib_truth)
            if r_c < TARGET_RECALL + CALIB_MARGIN:
                continue
            ms_c, _ = run_c.timed(ef, max(3, args.steps // 2), 2, o)
            if best is None or ms_c < best[0]:
                best = (ms_c, ef, budget, o, r_c)
    if best is not None:
        _, ef, budget, o, r_c = best
        run.launch(ef, o)
        r = recall(run.results(), truth)
        ms, _ = run.timed(ef, args.steps, args.warmup, o)
        budget_leg = {"ef": ef, "max_expansions": budget, "calib_recall": round(r_c, 4), "recall": round(r, 4),
                      "ms": round(ms, 4), "qps": round(10_000 / ms * 1000, 1),
                      "speedup_vs_fixed_ef": round(at[0]["ms"] / ms, 3) if at else None}
        log(f"10M budget: ef {ef}, at most {budget} expansions (held-out recall {r_c:.4f}): recall {r:.4f}, "
            f"{ms:.3f} ms, {budget_leg['qps']:,.0f} QPS")
    del run_c
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
            "qps_at_recall_0.95": at[0] if at else None, "budget_at_recall_0.95": budget_leg,
            "serving": {"queries": nq, "one_launch_ms": round(ms_one, 2), "tuned_graph_ms": round(ms_graph, 2),
                        "tuned_graph_qps": round(nq / ms_graph * 1000, 1), "identical_to_one_launch": identical}}


def e4m3_per_vector(v):
    """FP8-e4m3 with a per-vector scale as the GPU quantizes (scale max|v| / 448, round to
    nearest even, saturating): the dequantized rows."""
    v = np.asarray(v, dtype=np.float32)
    scale = np.abs(v).max(axis=1, keepdims=True) / np.float32(448.0)
    scale[scale == 0] = 1.0
    y = (v / scale).astype(np.float64)
    e = np.floor(np.log2(np.maximum(np.abs(y), 2.0 ** -20)))
    step = 2.0 ** (np.maximum(e, -6.0) - 3.0)  # 3 mantissa bits; subnormal step 2^-9
    q = np.clip(np.round(y / step) * step, -448.0, 448.0)  # numpy rounds halves to even
    return (q * scale).astype(np.float32)


def fp8_bound_study(base, queries, metric, ks=(1, 10, 100), sample=200):
    """Why the FP8 shortlist does not certify: over a sample of queries, the certification
    bound E (the largest |FP8 GEMM distance - exact distance| any row may have, from the
    quantization errors by Cauchy-Schwarz, the leading terms of the re-rank's test) against the
    room there is (the (S+1)-th smallest exact distance minus the k-th, S = 8k), and the
    fractions of queries certified: with the global bound (the smallest FP8 distance outside the
    S smallest, minus E, above the k-th exact distance) and with a per-row bound (every row
    outside the S smallest, its FP8 distance minus its own row's bound above the k-th)."""
    x = np.asarray(base, dtype=np.float32)
    q = np.ascontiguousarray(queries[:sample], dtype=np.float32)
    xq, qq = e4m3_per_vector(x), e4m3_per_vector(q)
    x_err = np.linalg.norm(xq - x, axis=1)
    x_norm = np.linalg.norm(x, axis=1)
    q_err = np.linalg.norm(qq - q, axis=1)
    q_qnorm = np.linalg.norm(qq, axis=1)
    scale = 2.0 if metric == 0 else 1.0  # L2 distances carry -2 q.x
    e_global = scale * (q_qnorm * x_err.max() + q_err * x_norm.max())
    qn2 = np.einsum("ij,ij->i", q, q)
    margins = {k: [] for k in ks}
    cert_global = {k: 0 for k in ks}
    cert_row = {k: 0 for k in ks}
    for c0 in range(0, len(q), 50):  # distances of 50 queries at a time (matrix products)
        dots = x @ q[c0:c0 + 50].T
        adots = xq @ qq[c0:c0 + 50].T
        for j in range(dots.shape[1]):
            i = c0 + j
            if metric == 0:
                exact = (x_norm ** 2 + qn2[i]) - 2 * dots[:, j]
                approx = (x_norm ** 2 + qn2[i]) - 2 * adots[:, j]
            else:
                exact = 1 - dots[:, j]
                approx = 1 - adots[:, j]
            row_bound = scale * (q_qnorm[i] * x_err + q_err[i] * x_norm)
            for k in ks:
                S = 8 * k
                part = np.partition(exact, (k - 1, S))
                dk, ds1 = float(part[k - 1]), float(part[S])
                margins[k].append(ds1 - dk)
                # rows outside the S smallest FP8 distances; L, the smallest of those, must clear
                # the k-th exact distance by the bound
                outside = np.argpartition(approx, S)[S:]
                cert_global[k] += float(approx[outside].min()) - e_global[i] > dk
                cert_row[k] += bool(((approx[outside] - row_bound[outside]) > dk).all())
    out = {"sample": len(q), "per_k": {}}
    for k in ks:
        m = np.array(margins[k])
        out["per_k"][f"k{k}"] = {
            "bound_median": round(float(np.median(e_global)), 4),
            "margin_median": round(float(np.median(m)), 4),
            "margin_p90": round(float(np.percentile(m, 90)), 4),
            "certified_with_global_bound": round(float(cert_global[k]) / len(q), 4),
            "certified_with_per_row_bounds": round(float(cert_row[k]) / len(q), 4)}
    return out


def bruteforce_leg(args, nat, bf, queries, metric, threads):
    """BruteForceIndex::search on the GPU over all (prepared) queries at k in {1, 10, 100}: the
    exact scan (the reference's result), then the FP32, BF16 and FP8 MFMA shortlists with exact
    re-rank, each checked byte for byte against the exact scan; the float64 pass on the GPU
    (the largest |float32 - float64| over the returned neighbours, and the one-ulp tie re-rank's
    agreement with the float64 oracle on a sample); plus the host's exact scan (the oracle's
    restatement of the reference, on all host threads) over a bounded sample of the queries."""
    from oracle import oracle as o
    base = bf.vectors()  # as stored: normalized under cosine
    nq, (n, d) = len(queries), base.shape
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(np.ascontiguousarray(queries, dtype=np.float32), ctx)
    leg = {}
    flops = 2.0 * nq * n * d
    sample = 200
    # the BF16 GEMM shape by measurement (one or two 32-query blocks per wave): the tuner keeps the
    # faster for the timings below
    t = bf.tune(queries, 10, reps=3)
    leg["bf16_gemm_shapes"] = {"ms_one_block_per_wave": round(t["ms"][0], 3),
                               "ms_two_blocks_per_wave": round(t["ms"][1], 3), "chosen": t["shape"]}
    log(f"BF16 GEMM shapes (k=10): one query block per wave {t['ms'][0]:.3f} ms, two {t['ms'][1]:.3f} ms; "
        f"kept shape {t['shape']}")
    for k in (1, 10, 100):
        bout = nat.DeviceArray((nq, k), nat.NEIGHBOR_DTYPE, ctx)
        bcnt = nat.DeviceArray(nq, np.uint32, ctx)
        row, exact = {}, None

        def call(prec_code, st):
            nat.check(nat.lib.vdx_gpu_bf_search_device(bf.gpu_handle(), dq.ptr, nq, k, prec_code, bout.ptr, k,
                                                       bcnt.ptr, ctypes.byref(st)))

        for prec in ("exact", "fp32", "bf16", "fp8"):
            st = nat.BfStats()
            times = []
            for _ in range(1 if prec in ("exact", "fp8") else 3):
                nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
                t0 = time.perf_counter()
                call(nat.PRECISIONS[prec], st)
                times.append((time.perf_counter() - t0) * 1000)
            res = bout.numpy()
            ms = min(times)
            if prec == "exact":
                exact = res
                row["exact"] = {"ms": round(ms, 2), "qps": round(nq / ms * 1000, 1)}
                continue
            peak = FP32_MFMA_TFLOPS if prec == "fp32" else (BF16_MFMA_TFLOPS if prec == "bf16" else FP8_MFMA_TFLOPS)
            row[prec] = {"ms": round(ms, 2), "qps": round(nq / ms * 1000, 1),
                         "identical_to_exact": bool(res.tobytes() == exact.tobytes()),
                         "exact_fallbacks": int(st.exact_fallback), "shortlist": int(st.shortlist),
                         "end_to_end_tflops": round(flops / ms / 1e9, 1),
                         "frac_of_mfma_peak": round(flops / ms / 1e9 / peak, 4)}
        row["bf16_speedup_vs_fp32"] = round(row["fp32"]["ms"] / row["bf16"]["ms"], 2)
        # the float64 pass (GPU): the error of the returned float32 distances, then the re-rank
        st = nat.BfStats()
        call(nat.PRECISIONS["bf16"] | nat.PREC_FP64_ERROR, st)
        same = bool(bout.numpy().tobytes() == exact.tobytes())
        row["fp64"] = {"max_abs_err": float(st.max_err_fp64), "max_rel_err": float(st.max_rel_err_fp64),
                       "results_unchanged": same}
        st = nat.BfStats()
        t0 = time.perf_counter()
        call(nat.PRECISIONS["bf16"] | nat.PREC_FP64_TIES, st)
        ties_ms = (time.perf_counter() - t0) * 1000
        tied = bout.numpy()
        s64 = 50  # the float64 oracle is a plain double loop: a small sample
        ids64, _ = o.bruteforce_f64(base, np.ascontiguousarray(queries[:s64]), k, metric=metric, threads=threads)
        row["fp64"].update({"ties_ms": round(ties_ms, 2), "tie_candidates": int(st.fp64_ties),
                            "queries_reordered": int(st.fp64_reordered), "window": int(st.fp64_window),
                            "fp64_oracle_sample": s64,
                            "sample_ids_equal_fp64_oracle_fp32": float((exact["id"][:s64] == ids64).all(axis=1).mean()),
                            "sample_ids_equal_fp64_oracle_reranked": float((tied["id"][:s64] == ids64).all(axis=1).mean())})
        log(f"brute force k={k} ({n:,} x {d}): exact {row['exact']['ms']:.1f} ms, fp32 {row['fp32']['ms']:.2f}, "
            f"bf16 {row['bf16']['ms']:.2f}, fp8 {row['fp8']['ms']:.1f} ms (fallbacks {row['fp8']['exact_fallbacks']}); "
            f"identical {[row[p]['identical_to_exact'] for p in ('fp32', 'bf16', 'fp8')]}; max |fp32 - fp64| "
            f"{st.max_err_fp64:.3g}")
        leg[f"k{k}"] = row
        del bout, bcnt
    t0 = time.perf_counter()
    leg["fp8_bounds"] = fp8_bound_study(base, queries, metric)
    log(f"FP8 certification study ({time.perf_counter() - t0:.1f} s): " + "; ".join(
        f"{kk}: bound {v['bound_median']:.4g} vs margin {v['margin_median']:.4g} (p90 {v['margin_p90']:.4g}), "
        f"certified {v['certified_with_global_bound']:.0%} global / {v['certified_with_per_row_bounds']:.0%} per row"
        for kk, v in leg["fp8_bounds"]["per_k"].items()))
    if not args.no_cpu:
        t0 = time.perf_counter()
        host = o.bruteforce_f32(base, np.ascontiguousarray(queries[:sample]), K, metric=metric, threads=threads)
        secs = time.perf_counter() - t0
        gpu10 = leg["k10"]
        leg["cpu_exact"] = {"qps": round(sample / secs, 1), "cores": threads, "kind": "port",
                            "sample": f"{sample} of the {nq:,} queries at k=10, oracle exact scan (FP32, AVX2 order)"}
        log(f"host exact scan: {sample / secs:,.0f} QPS on {threads} threads (sample of {sample}); GPU/host "
            f"speed-up of the BF16 path {gpu10['bf16']['qps'] / (sample / secs):,.0f}x")
    del dq
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


def run_budget(args, nat, index, queries, truth, calib, calib_truth, fixed_best, headline):
    """Fixed ef with an expansion budget (vdx_search_opts.max_expansions): every query stops after
    at most B layer-0 expansions. A batch takes as long as its slowest queries, and a budget cuts
    exactly those, so it can gain more time than the expansions it saves. (ef, B) is chosen on the
    held-out queries (recall >= target + CALIB_MARGIN, fastest there), then measured on the
    evaluation queries."""
    rc, re_ = Runner(nat, index, calib), Runner(nat, index, queries)
    table, best = [], None
    for budget in map(int, args.budgets.split(",")):
        for ef in map(int, args.budget_efs.split(",")):
            o = nat.SearchOpts()
            o.max_expansions = budget
            rc.launch(ef, o)
            r_c = recall(rc.results(), calib_truth)
            if r_c < TARGET_RECALL + CALIB_MARGIN:
                continue
            ms_c, _ = rc.timed(ef, max(3, args.steps // 2), 2, o)
            table.append({"ef": ef, "max_expansions": budget, "calib_recall": round(r_c, 4), "calib_ms": round(ms_c, 4)})
            if best is None or ms_c < best[0]:
                best = (ms_c, ef, budget, o, r_c)
    del rc
    if best is None:
        return {"table": table}, headline
    _, ef, budget, o, r_c = best
    re_.launch(ef, o)
    rec = recall(re_.results(), truth)
    ms, _ = re_.timed(ef, args.steps, args.warmup, o)
    counts = re_.counters(ef, o)
    leg = {"ef": ef, "max_expansions": budget, "calib_recall": round(r_c, 4), "eval_recall": round(rec, 4),
           "ms": round(ms, 4), "qps": round(len(queries) / ms * 1000, 1),
           "expansions": round(counts[0] / len(queries), 1),
           "stopped_by_budget": round(counts[4] / len(queries), 4),
           "fixed_ef_at_target": {k: fixed_best[k] for k in ("ef", "recall", "qps")},
           "speedup_vs_fixed_ef": round(len(queries) / ms * 1000 / fixed_best["qps"], 3),
           "table": table}
    log(f"budget: ef {ef}, at most {budget} expansions (held-out recall {r_c:.4f}): eval recall {rec:.4f}, "
        f"{ms:.3f} ms, {leg['qps']:,.0f} QPS ({leg['speedup_vs_fixed_ef']}x fixed ef {fixed_best['ef']}); "
        f"{leg['expansions']} expansions, {leg['stopped_by_budget'] * 100:.1f}% of queries stopped by the budget")
    if rec >= TARGET_RECALL and leg["qps"] > headline["qps"]:
        headline = {"mode": "budget", "ef": ef, "opts": o, "qps": leg["qps"], "recall": rec}
    del re_
    return leg, headline


def run_learned(args, vdx, nat, datasets, learned, index, base, queries, truth, mix, threads,
                fixed, fixed_best, headline, budget_leg=None):
    """Traces, training, calibration and evaluation of the learned stop on both bases, and the
    oracle-stop ceiling of each.

    Recall target 0.95 on both bases. Cosine 1.2M x 100: a fixed-ef sweep up to 2048 on the
    evaluation queries (above 512 the LDS beam kernel); the cap is the smallest ef whose
    held-out recall reaches the target. Traces: training queries of each base at its trace caps
    (L2: ef 64 and 128, half the queries each; cosine: the cap, on as many queries as 400 MB of
    trace holds), labelled by brute-force top-10 hits; one model for both (20 epochs, Adam
    lr 1e-3, batch 4096, up to --max-train-steps sampled steps). Calibration on held-out queries
    (never the evaluation set): per ef cap and consult interval (every 4th or 8th expansion) the
    threshold with the fewest expansions reaching target + 0.002, then the setting with the
    highest calibration QPS.

    Ceiling: the evaluation queries' traced searches at each cap, stopped by an oracle at the
    first expansion whose top-10 equals the final top-10, and by a greedy recall-budgeted oracle
    (the fewest expansions whose mean recall still reaches 0.95), against the expansions fixed ef
    needs for 0.95: the most any stopping rule inside these searches could save."""
    t_all = time.perf_counter()
    cos_gen, cos_base, cos_index = cosine_setup(args, vdx, datasets, threads)
    ntr = args.train_queries
    sets = {
        "l2": {"index": index, "base": base, "metric": vdx.Metric.L2, "trace_efs": (64, 128),
               "caps": (40, 48, 64, 96, 128),
               "train": mix.sample(ntr, 3), "calib": mix.sample(args.nq, 4), "eval": queries},
        "cos": {"index": cos_index, "base": cos_base, "metric": vdx.Metric.Cosine,
                "train": prepared(nat, cos_gen.sample(ntr, 3), True),
                "calib": prepared(nat, cos_gen.sample(args.nq, 4), True),
                "eval": prepared(nat, cos_gen.sample(args.nq, 2), True)},
    }
    for name, s in sets.items():
        allq = np.concatenate([s["train"], s["calib"], s["eval"]])
        t = time.perf_counter()
        tr, _ = ground_truth(vdx, s["base"], s["metric"], allq)
        s["truth_train"] = tr[:ntr]
        s["truth_calib"] = tr[ntr:ntr + args.nq]
        s["truth_eval"] = truth if name == "l2" else tr[ntr + args.nq:]
        log(f"learned/{name}: ground truth for {len(allq)} queries in {time.perf_counter() - t:.1f} s")

    # the cosine base's fixed-ef sweep; its cap for recall 0.95 from the held-out queries
    s = sets["cos"]
    rc, re_ = Runner(nat, s["index"], s["calib"]), Runner(nat, s["index"], s["eval"])
    sweep_cos, fixed_cos = [], None
    for ef in (64, 128, 256, 320, 384, 448, 512, 640, 768, 1024, 1536, 2048):
        re_.launch(ef)
        r_e = recall(re_.results(), s["truth_eval"])
        rc.launch(ef)
        r_c = recall(rc.results(), s["truth_calib"])
        ms, form = timed_best_form(re_, ef, max(3, args.steps // 2), 2)
        sweep_cos.append({"ef": ef, "recall": round(r_e, 4), "calib_recall": round(r_c, 4), "ms": round(ms, 4),
                          "qps": round(len(s["eval"]) / ms * 1000, 1), "visited_form": form})
        log(f"cos fixed ef {ef}: recall {r_e:.4f} (held-out {r_c:.4f}), {sweep_cos[-1]['qps']:,.0f} QPS")
        if fixed_cos is None and r_c >= TARGET_RECALL:
            fixed_cos = sweep_cos[-1]
    del rc, re_
    cap_cos = fixed_cos["ef"] if fixed_cos else 2048
    s["trace_efs"] = (cap_cos,)
    s["caps"] = tuple(sorted({cap_cos, int(cap_cos * 1.5), 2 * cap_cos} & set(range(1, 8193))))

    # traces and training
    feats, labels = [], []
    for name, s in sets.items():
        efs = s["trace_efs"]
        per = ntr // len(efs)
        for part, ef in enumerate(efs):
            max_steps = 2 * ef
            nq_tr = min(per, max(1000, (400 << 20) // (max_steps * 16 * 4)))
            lo = part * per
            t = time.perf_counter()
            f, lab, steps, final = learned.collect_traces(s["index"], s["train"][lo:lo + nq_tr],
                                                          s["truth_train"]["id"][lo:lo + nq_tr], k=K, ef=ef,
                                                          max_steps=max_steps)
            log(f"learned/{name}: traces at ef {ef} over {nq_tr} queries: {len(f):,} steps ({steps.mean():.1f} "
                f"expansions/query, label mean {lab.mean():.3f}) in {time.perf_counter() - t:.1f} s")
            feats.append(f)
            labels.append(lab)
    f = np.concatenate(feats)
    lab = np.concatenate(labels)
    if len(f) > args.max_train_steps:
        pick = np.random.default_rng(0).choice(len(f), args.max_train_steps, replace=False)
        f, lab = f[pick], lab[pick]
    n_train = len(f)
    t = time.perf_counter()
    model, losses = learned.train_stop_model(f, lab, epochs=20, batch=4096, lr=1e-3, seed=0)
    train_s = time.perf_counter() - t
    log(f"learned: trained on {len(f):,} steps, 20 epochs in {train_s:.1f} s; loss "
        f"{losses[0]:.4f} -> {losses[-1]:.4f}")
    del feats, labels, f, lab

    results = {}
    fixed_eval = {"l2": {"ef": fixed_best["ef"], "recall": fixed_best["recall"], "qps": fixed_best["qps"]},
                  "cos": fixed_cos}
    for name, s in sets.items():
        target = TARGET_RECALL
        rc = Runner(nat, s["index"], s["calib"])
        re_ = Runner(nat, s["index"], s["eval"])
        # the ceiling, from the evaluation queries' traced searches
        ceiling = []
        fixed_qps = fixed_eval[name]["qps"] if fixed_eval[name] else None
        for cap in ((36, 48, 64, 128) if name == "l2" else (cap_cos,)):
            hits, steps, final = learned.trace_hits(s["index"], s["eval"], s["truth_eval"]["id"], K, cap,
                                                    max_steps=2 * cap)
            c = learned.stop_ceiling(hits, steps, final, target, K, per_query=True)
            c["cap"] = cap
            log(f"learned/{name} ceiling at cap {cap}: fixed {c['fixed_expansions']} expansions (recall "
                f"{c['fixed_recall']}), oracle stop {c['oracle_stop_expansions']}, budgeted to {target}: "
                f"{c.get('budgeted_expansions')}")
            # The ceiling in time: the same searches run with the oracle's per-query stops as
            # per-query expansion budgets (a stop at 0 expansions runs 1), timed like the others.
            for key in ("oracle_stops", "budgeted_stops"):
                stops = c.pop(key, None)
                if stops is None:
                    continue
                db = nat.DeviceArray.from_numpy(np.maximum(stops, 1).astype(np.uint32), re_.ctx)
                ob = nat.SearchOpts()
                ob.max_expansions_per_query = db.ptr
                re_.launch(cap, ob)
                rec_b = recall(re_.results(), s["truth_eval"])
                ms_b, _ = timed_best_form(re_, cap, max(3, args.steps // 2), 2, ob)
                t = {"recall": round(rec_b, 4), "ms": round(ms_b, 4), "qps": round(len(s["eval"]) / ms_b * 1000, 1)}
                if fixed_qps:
                    t["speedup_vs_fixed_ef"] = round(t["qps"] / fixed_qps, 3)
                c[key.replace("_stops", "_stop_time")] = t
                log(f"learned/{name} ceiling at cap {cap}, {key} as per-query budgets: recall {rec_b:.4f}, "
                    f"{ms_b:.3f} ms, {t['qps']:,.0f} QPS ({t.get('speedup_vs_fixed_ef')}x fixed ef)")
                del db
            ceiling.append(c)
        best = None
        # expansion budgets tried with the model: none, and (L2) the budget leg's and a looser one
        budgets = [0]
        if name == "l2" and budget_leg and budget_leg.get("max_expansions"):
            budgets += [budget_leg["max_expansions"], budget_leg["max_expansions"] + 8]
        for ef_cap in s["caps"]:
            for every in (4, 8):  # consult the model every 4th or 8th expansion
                for budget in budgets:
                    thr, rec_c, work, table = learned.calibrate(s["index"], model, s["calib"],
                                                                s["truth_calib"]["id"], K, ef_cap,
                                                                target + CALIB_MARGIN, every=every, grid=THRESHOLDS,
                                                                max_expansions=budget)
                    if rec_c < target + CALIB_MARGIN:
                        log(f"learned/{name} ef cap {ef_cap}, every {every}, budget {budget}: no threshold "
                            f"reaches {target + CALIB_MARGIN:.4f} (best {rec_c:.4f})")
                        continue
                    opts = learned.learned_opts(model, thr, 1, every=every, max_expansions=budget)
                    ms_c, form = timed_best_form(rc, ef_cap, max(3, args.steps // 2), 2, opts)
                    opts.visited_form = form
                    log(f"learned/{name} ef cap {ef_cap}, every {every}, budget {budget}: threshold {thr}, "
                        f"calib recall {rec_c:.4f}, {work:.1f} expansions, "
                        f"{len(s['calib']) / ms_c * 1000:,.0f} QPS on calibration")
                    if best is None or ms_c < best["_ms_c"]:
                        best = {"ef_cap": ef_cap, "every": every, "threshold": thr, "max_expansions": budget,
                                "calib_recall": round(rec_c, 4), "_ms_c": ms_c, "_opts": opts}
        fe = fixed_eval[name]
        if fe is not None:
            # the fixed search's expansions at its operating point, on the evaluation queries
            fe = dict(fe)
            fe["expansions"] = round(re_.counters(fe["ef"])[0] / len(s["eval"]), 1)
        if best is not None:
            opts = best["_opts"]
            re_.launch(best["ef_cap"], opts)
            rec_e = recall(re_.results(), s["truth_eval"])
            ms, _ = re_.timed(best["ef_cap"], args.steps, args.warmup, opts)
            exp = re_.counters(best["ef_cap"], opts)[0] / len(s["eval"])
            best.update({"target": target, "eval_recall": round(rec_e, 4), "ms": round(ms, 4),
                         "qps": round(len(s["eval"]) / ms * 1000, 1), "expansions": round(exp, 1),
                         "fixed_ef_at_target": fe})
            if fe is not None:
                best["speedup_vs_fixed_ef"] = round(best["qps"] / fe["qps"], 3)
            if name == "l2" and budget_leg and budget_leg.get("qps"):
                best["speedup_vs_budget"] = round(best["qps"] / budget_leg["qps"], 3)
            log(f"learned/{name}: eval recall {rec_e:.4f} (target {target}), {ms:.3f} ms, "
                f"{best['qps']:,.0f} QPS, {exp:.1f} expansions; fixed ef at target: {fe}")
        if fe is not None:
            floor = min((c["budgeted_expansions"] for c in ceiling if c.get("budgeted_expansions")), default=None)
            if floor:
                for c in ceiling:
                    c["fixed_expansions_at_target"] = fe["expansions"]
                results[name + "_ceiling_speedup"] = round(fe["expansions"] / floor, 3)
        results[name] = best
        results[name + "_ceiling"] = ceiling
        del rc, re_
    l2 = results["l2"]
    leg = {
        "train_steps": n_train,
        "train_s": round(train_s, 1),
        "loss_first_last": [round(float(losses[0]), 4), round(float(losses[-1]), 4)],
        "l2": {k: v for k, v in (l2 or {}).items() if not k.startswith("_")},
        "cos": {k: v for k, v in (results["cos"] or {}).items() if not k.startswith("_")},
        "oracle_stop_ceiling": {"l2": results["l2_ceiling"], "cos": results["cos_ceiling"],
                                "l2_max_speedup_in_expansions": results.get("l2_ceiling_speedup"),
                                "cos_max_speedup_in_expansions": results.get("cos_ceiling_speedup")},
        "cos_fixed_sweep": sweep_cos,
        "total_s": round(time.perf_counter() - t_all, 1),
    }
    if l2 is not None and l2["eval_recall"] >= TARGET_RECALL and l2["qps"] > headline["qps"]:
        headline = {"mode": "learned", "ef": l2["ef_cap"], "opts": l2["_opts"], "qps": l2["qps"],
                    "recall": l2["eval_recall"], "model": model}  # the model must stay alive
    return leg, headline


if __name__ == "__main__":
    sys.exit(main())
