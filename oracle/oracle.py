"""ctypes wrapper of the C oracle (vdx_oracle.c) plus numpy restatements of vecdex's recall
functions and result fingerprints.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module. It is the checker, never the thing measured or shipped.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libvdx_oracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


if not LIB_PATH.exists():
    build()

_lib = ctypes.CDLL(str(LIB_PATH))
_f, _u32, _u64, _int, _p = (ctypes.c_float, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                            ctypes.c_void_p)

for _name in ("vdxo_l2_avx2order", "vdxo_ip_avx2order", "vdxo_l2_scalar", "vdxo_ip_scalar",
              "vdxo_l2_fast", "vdxo_ip_fast"):
    getattr(_lib, _name).restype = _f
    getattr(_lib, _name).argtypes = [_p, _p, ctypes.c_size_t]
_lib.vdxo_has_avx2.restype = _int
_lib.vdxo_normalize.restype = None
_lib.vdxo_normalize.argtypes = [_p, ctypes.c_size_t]
_lib.vdxo_bruteforce_f64.restype = _u64
_lib.vdxo_bruteforce_f64.argtypes = [_p, _u64, _u64, _u32, _p, _u64, _u64, _p, _p, _int]
_lib.vdxo_bruteforce_f32.restype = _u64
_lib.vdxo_bruteforce_f32.argtypes = [_p, _u64, _u64, _u32, _p, _u64, _u64, _p, _int]
_lib.vdxo_xxh64.restype = _u64
_lib.vdxo_xxh64.argtypes = [_p, ctypes.c_size_t, _u64]


class _Graph(ctypes.Structure):
    _fields_ = [("metric", _u32), ("dim", _u64), ("n", _u64), ("M", _u64), ("max_level", _u64),
                ("entry_point", _u64), ("levels", _p), ("layer0", _p), ("upper", _p),
                ("vectors", _p)]


_lib.vdxo_hnsw_search_batch.restype = _int
_lib.vdxo_hnsw_search_batch.argtypes = [ctypes.POINTER(_Graph), _p, _u64, _u64, _u64, _p, _u64,
                                        _p, _int, _int]
_lib.vdxo_hnsw_search_batch_bounded.restype = _int
_lib.vdxo_hnsw_search_batch_bounded.argtypes = [ctypes.POINTER(_Graph), _p, _u64, _u64, _u64, _u64, _p, _u64,
                                                _p, _p, _int, _int]
_lib.vdxo_hnsw_search_batch_from.restype = _int
_lib.vdxo_hnsw_search_batch_from.argtypes = [ctypes.POINTER(_Graph), _p, _p, _u64, _u64, _u64, _u64, _p,
                                             _u64, _p, _int, _int]

NEIGHBOR = np.dtype([("distance", np.float32), ("id", np.uint32)])


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def distance(kind: str, a, b) -> float:
    """kind: l2_avx2order, ip_avx2order, l2_scalar, ip_scalar, l2_fast, ip_fast."""
    a, b = _c(a, np.float32), _c(b, np.float32)
    return float(getattr(_lib, "vdxo_" + kind)(a.ctypes.data, b.ctypes.data, a.size))


def has_avx2() -> bool:
    return bool(_lib.vdxo_has_avx2())


def normalize(v) -> np.ndarray:
    out = np.array(v, dtype=np.float32, copy=True)
    for row in out.reshape(-1, out.shape[-1]):
        _lib.vdxo_normalize(row.ctypes.data, row.size)
    return out


def bruteforce_f64(base, queries, k: int, metric: int = 0, threads: int = 0):
    """(ids uint32[nq, keep], distances float64[nq, keep]), ties by smaller id."""
    base, queries = _c(base, np.float32), _c(queries, np.float32)
    n, d = base.shape
    nq = queries.shape[0]
    keep = min(k, n)
    ids = np.zeros((nq, max(k, 1)), dtype=np.uint32)
    dist = np.zeros((nq, max(k, 1)), dtype=np.float64)
    _lib.vdxo_bruteforce_f64(base.ctypes.data, n, d, metric, queries.ctypes.data, nq, k,
                             ids.ctypes.data, dist.ctypes.data, threads)
    return ids[:, :keep], dist[:, :keep]


def bruteforce_f32(base, queries, k: int, metric: int = 0, threads: int = 0) -> np.ndarray:
    """BruteForceIndex::search per query over stored (already normalized) vectors:
    structured (distance, id) array [nq, min(k, n)]."""
    base, queries = _c(base, np.float32), _c(queries, np.float32)
    n, d = base.shape
    nq = queries.shape[0]
    keep = min(k, n)
    out = np.zeros((nq, max(k, 1)), dtype=NEIGHBOR)
    _lib.vdxo_bruteforce_f32(base.ctypes.data, n, d, metric, queries.ctypes.data, nq, k,
                             out.ctypes.data, threads)
    return out[:, :keep]


def hnsw_search(graph: dict, queries, k: int, ef: int, threads: int = 0, fast: bool = True,
                entries=None, layer: int = 0, max_expansions: int | None = None):
    """HnswIndex::search for every row of prepared queries (normalized under cosine) over a
    graph given as FORMAT v1 sections (dict of numpy arrays and header fields).
    layer: the layer the beam searches (0 for queries; insert searches each of its layers).
    entries: per query, an entry on that layer that replaces the greedy descent (insert's
    searches start from the layer above's result, src/hnsw.cpp:324-335), or 0xFFFFFFFF.
    max_expansions (layer 0, no entries; not vecdex's): the expansion budget, 0 for none; then
    the result also carries each query's expansions taken.
    Returns (structured [nq, k] array, counts[nq]) [, expansions[nq]]."""
    queries = _c(queries, np.float32)
    keep = [graph[name] for name in ("levels", "layer0", "upper", "vectors")]
    g = _Graph(graph["metric"], graph["dim"], graph["n"], graph["M"], graph["max_level"],
               graph["entry_point"], *(a.ctypes.data if a.size else None for a in keep))
    nq = queries.shape[0]
    out = np.zeros((nq, k), dtype=NEIGHBOR)
    count = np.zeros(nq, dtype=np.uint32)
    if max_expansions is not None:
        assert entries is None and layer == 0
        taken = np.zeros(nq, dtype=np.uint32)
        _lib.vdxo_hnsw_search_batch_bounded(ctypes.byref(g), queries.ctypes.data, nq, k, ef, int(max_expansions),
                                            out.ctypes.data, k, count.ctypes.data, taken.ctypes.data, threads,
                                            1 if fast else 0)
        del keep
        return out, count, taken
    ent = None if entries is None else _c(entries, np.uint32)
    _lib.vdxo_hnsw_search_batch_from(ctypes.byref(g), queries.ctypes.data,
                                     None if ent is None else ent.ctypes.data, layer, nq, k, ef, out.ctypes.data,
                                     k, count.ctypes.data, threads, 1 if fast else 0)
    del keep
    return out, count


def xxh64(data: bytes, seed: int = 0) -> int:
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    return int(_lib.vdxo_xxh64(buf, len(data), seed))


def fingerprint(results) -> int:
    """FNV-1a over ids and distance bits (tests/hnsw_parallel_test.cpp:91-105). `results` is a
    list of (ids, distances) per query."""
    h = 14695981039346656037
    mask = (1 << 64) - 1
    for ids, dists in results:
        bits = np.asarray(dists, dtype=np.float32).view(np.uint32)
        for i, b in zip(np.asarray(ids, dtype=np.uint32), bits):
            for word in (int(i), int(b)):
                for shift in range(0, 32, 8):
                    h = ((h ^ ((word >> shift) & 0xFF)) * 1099511628211) & mask
    return h


def recall_at_k(found_ids, truth_ids, k: int) -> float:
    """recall_at_k (src/recall.cpp:11-34), averaged over queries."""
    f = np.asarray(found_ids)[:, :k]
    t = np.asarray(truth_ids)[:, :k]
    return float((t[:, :, None] == f[:, None, :]).any(axis=2).mean())


def threshold_recall_at_k(found_dists, kth_true, k: int, tolerance: float = 1e-5) -> float:
    """threshold_recall_at_k (src/recall.cpp:36-51), averaged over queries: results within the
    k-th true distance plus tolerance * max(|d_k|, 1)."""
    kth = np.asarray(kth_true, dtype=np.float32)
    thr = kth + np.float32(tolerance) * np.maximum(np.abs(kth), np.float32(1.0))
    d = np.asarray(found_dists, dtype=np.float32)[:, :k]
    return float((d <= thr[:, None]).sum(axis=1).mean() / k)
