"""The evaluation sweep of vecdex_eval (bench/hnsw_eval.cpp:20-26, 415-527) on the GPU path.

    python -m hnsw_vector_search_amd.eval <dir> <name> [--M 16] [--ef-construction 200] [--k 10]
        [--ef 10,16,32,64,128,256,512] [--repeats 5] [--threads 1] [--save FILE] [--load FILE]
        [--device-resident]

Reads <dir>/<name>_base.fvecs, <name>_query.fvecs and <name>_groundtruth.ivecs (L2), builds the
graph on the host (or loads a FORMAT v1 file) and prints vecdex_eval's CSV,
`lib,dataset,n,dim,M,ef_construction,build_s,memory_mb,k,ef,recall,recall_dist,qps,threads`,
one row per ef, with lib = "vecdex-mi355x". QPS is the median of `repeats` timed passes over all
queries after one untimed warmup, as vecdex_eval times them: by default through the host entry
point (HnswIndex::search_batch: queries from host memory, results back to host memory); with
--device-resident, queries and results stay in HBM and passes are timed with HIP events. Each row
also prints vecdex_eval's FNV-1a results fingerprint to stderr, so runs can be compared with a
vecdex_eval sweep over the same graph.
"""

from __future__ import annotations

import argparse
import sys
import time

import numpy as np

USAGE = ("usage: python -m hnsw_vector_search_amd.eval <dir> <name> [--M 16] "
         "[--ef-construction 200] [--k 10] [--ef 10,16,32,64,128,256,512] [--repeats 5] "
         "[--threads 1] [--save FILE] [--load FILE] [--device-resident]")


def parse(argv):
    """vecdex_eval's options (bench/hnsw_eval.cpp:316-367): sizes are positive integers, --ef a
    non-empty comma-separated list. Returns None on a bad command line."""
    ap = argparse.ArgumentParser(add_help=False)
    ap.add_argument("dir")
    ap.add_argument("name")
    for flag, default in (("--M", 16), ("--ef-construction", 200), ("--k", 10), ("--repeats", 5),
                          ("--threads", 1)):
        ap.add_argument(flag, type=str, default=str(default))
    ap.add_argument("--ef", type=str, default="10,16,32,64,128,256,512")
    ap.add_argument("--save", type=str, default=None)
    ap.add_argument("--load", type=str, default=None)
    ap.add_argument("--device-resident", action="store_true")
    try:
        a = ap.parse_args(argv)
    except SystemExit:
        return None
    try:
        for name in ("M", "ef_construction", "k", "repeats", "threads"):
            v = getattr(a, name)
            if not v.isdigit() or int(v) == 0:
                return None
            setattr(a, name, int(v))
        parts = a.ef.split(",")
        if not parts or any(not p.isdigit() for p in parts):
            return None
        a.efs = [int(p) for p in parts]
    except ValueError:
        return None
    return a


def check_truth(truth: np.ndarray, queries: int, base_rows: int, k: int) -> None:
    """The ground truth covers every query with at least k ids naming base vectors
    (bench/hnsw_eval.cpp:415-428)."""
    if truth.shape[0] < queries or truth.shape[1] < k:
        raise RuntimeError(f"ground truth is {truth.shape[0]} x {truth.shape[1]}, need at least "
                           f"{queries} x {k}")
    head = truth[:queries, :k]
    bad = (head < 0) | (head >= base_rows)
    if bad.any():
        raise RuntimeError(f"ground truth names missing vector {int(head[bad][0])}")


def fingerprint(out: np.ndarray, counts: np.ndarray) -> int:
    """FNV-1a over each query's result count, then its ids and distance bits, byte by byte
    (bench/hnsw_eval.cpp:386-402)."""
    mask = (1 << 64) - 1
    prime = 1099511628211
    h = 14695981039346656037
    ids = out["id"].astype(np.uint32)
    bits = out["distance"].astype(np.float32).view(np.uint32)
    for q in range(len(counts)):
        c = int(counts[q])
        words = [c]
        for i in range(c):
            words.append(int(ids[q, i]))
            words.append(int(bits[q, i]))
        for w in words:
            for shift in (0, 8, 16, 24):
                h = ((h ^ ((w >> shift) & 0xFF)) * prime) & mask
    return h


def row(lib, dataset, n, dim, M, efc, build_s, memory_mb, k, ef, recall, recall_dist, qps, threads):
    return (f"{lib},{dataset},{n},{dim},{M},{efc},{build_s:.1f},{memory_mb:.0f},{k},{ef},"
            f"{recall:.4f},{recall_dist:.4f},{qps:.0f},{threads}")


def run(argv) -> int:
    a = parse(argv)
    if a is None:
        print(USAGE, file=sys.stderr)
        return 2
    from . import _native as nat
    from . import index as ix
    from . import recall as rc
    from .vecs import read_fvecs, read_ivecs

    prefix = f"{a.dir.rstrip('/')}/{a.name}"
    base = read_fvecs(prefix + "_base.fvecs")
    queries = read_fvecs(prefix + "_query.fvecs")
    truth = read_ivecs(prefix + "_groundtruth.ivecs")
    k = a.k
    if queries.shape[1] != base.shape[1]:
        raise RuntimeError("queries and base vectors have different dimensions")
    check_truth(truth, len(queries), len(base), k)
    print(f"loaded {len(base)} x {base.shape[1]} base vectors, {len(queries)} queries", file=sys.stderr)
    # the k-th true neighbour's distance, in the library's arithmetic (exact on SIFT's integers)
    kth = np.array([nat.lib.vdx_host_distance(0, nat.ptr(queries[q]), nat.ptr(base[truth[q, k - 1]]),
                                              base.shape[1]) for q in range(len(queries))], dtype=np.float32)
    build_s = float("nan")
    if a.load:
        t = time.perf_counter()
        index = ix.HnswIndex.load(a.load)
        if index.dim() != base.shape[1] or index.metric() != ix.Metric.L2:
            raise RuntimeError(f"{a.load} is not an L2 index of dimension {base.shape[1]}")
        M, efc = index.params()["M"], index.params()["ef_construction"]
        print(f"loaded {a.load} in {time.perf_counter() - t:.2f} s", file=sys.stderr)
    else:
        index = ix.HnswIndex(base.shape[1], ix.Metric.L2, M=a.M, ef_construction=a.ef_construction)
        t = time.perf_counter()
        index.add(base, num_threads=a.threads)
        build_s = time.perf_counter() - t
        M, efc = a.M, a.ef_construction
        print(f"built in {build_s:.1f} s on {a.threads} thread(s)", file=sys.stderr)
    if a.save:
        index.save(a.save)
    memory_mb = index.memory_bytes() / 1e6
    truth_k = np.ascontiguousarray(truth[:len(queries), :k])

    print("lib,dataset,n,dim,M,ef_construction,build_s,memory_mb,k,ef,recall,recall_dist,qps,threads")
    ctx = dq = dout = dcnt = None
    if a.device_resident:
        ctx = nat.Context.get(0)
        dq = nat.DeviceArray.from_numpy(queries, ctx)
        dout = nat.DeviceArray((len(queries), k), nat.NEIGHBOR_DTYPE, ctx)
        dcnt = nat.DeviceArray(len(queries), np.uint32, ctx)
    for ef in a.efs:
        if a.device_resident:
            def one_pass():
                nat.check(nat.lib.vdx_gpu_hnsw_search_device(index.gpu_handle(), dq.ptr, len(queries), None,
                                                             k, None, ef, k, dout.ptr, k, dcnt.ptr, None, None))
                nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
                return dout.numpy(), dcnt.numpy()
        else:
            def one_pass():
                return index.search_arrays(queries, k, ef)
        out, counts = one_pass()  # warmup
        times = []
        for _ in range(a.repeats):
            if a.device_resident:
                nat.check(nat.lib.vdx_gpu_event_record(ctx.handle, 4))
                nat.check(nat.lib.vdx_gpu_hnsw_search_device(index.gpu_handle(), dq.ptr, len(queries), None,
                                                             k, None, ef, k, dout.ptr, k, dcnt.ptr, None, None))
                nat.check(nat.lib.vdx_gpu_event_record(ctx.handle, 5))
                nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
                times.append(nat.elapsed_ms(ctx, 4, 5) / 1000.0)
            else:
                t = time.perf_counter()
                out, counts = one_pass()
                times.append(time.perf_counter() - t)
        if a.device_resident:
            out, counts = dout.numpy(), dcnt.numpy()
        median = sorted(times)[len(times) // 2]
        print(f"ef {ef}: results fingerprint {fingerprint(out, counts):016x}", file=sys.stderr)
        recall = rc.mean_recall_at_k(out, truth_k, k, counts=counts)
        recall_dist = rc.mean_threshold_recall_at_k(out, kth, k, counts=counts)
        print(row("vecdex-mi355x", a.name, len(base), base.shape[1], M, efc, build_s, memory_mb, k, ef,
                  recall, recall_dist, len(queries) / median, a.threads), flush=True)
    return 0


def main(argv=None) -> int:
    try:
        return run(sys.argv[1:] if argv is None else argv)
    except Exception as e:  # unreadable dataset, bad parameters, ...
        print(f"vecdex_eval: {e}", file=sys.stderr)
        return 1


if __name__ == "__main__":
    sys.exit(main())
