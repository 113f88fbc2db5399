"""Recall measurement: vecdex's recall_at_k and threshold_recall_at_k (src/recall.cpp:11-51)
through the library's C ABI, plus means over result matrices for batched evaluation.

`found` is a sequence of Neighbor (or a structured array of NEIGHBOR_DTYPE); `truth` holds the
true neighbour ids best first. Bad arguments raise ValueError with the reference's messages
(std::invalid_argument in vecdex).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat


def _neighbors(found) -> np.ndarray:
    if isinstance(found, np.ndarray) and found.dtype == nat.NEIGHBOR_DTYPE:
        return np.ascontiguousarray(found)
    out = np.zeros(len(found), dtype=nat.NEIGHBOR_DTYPE)
    for i, n in enumerate(found):
        out[i] = (n.distance, n.id) if hasattr(n, "distance") else (n[0], n[1])
    return out


def recall_at_k(found, truth, k: int) -> float:
    """|first k of found ∩ first k of truth| / k (src/recall.cpp:11-34)."""
    f = _neighbors(found)
    t = np.ascontiguousarray(truth, dtype=np.int32)
    out = ctypes.c_double()
    nat.check(nat.lib.vdx_recall_at_k(nat.ptr(f) if f.size else None, f.size,
                                      nat.ptr(t) if t.size else None, t.size, int(k),
                                      ctypes.byref(out)))
    return out.value


def threshold_recall_at_k(found, kth_true_distance: float, k: int, tolerance: float = 1e-5) -> float:
    """Fraction of the first k found within kth_true + tolerance * max(|kth_true|, 1)
    (src/recall.cpp:36-51)."""
    f = _neighbors(found)
    out = ctypes.c_double()
    nat.check(nat.lib.vdx_threshold_recall_at_k(nat.ptr(f) if f.size else None, f.size,
                                                float(kth_true_distance), int(k), float(tolerance),
                                                ctypes.byref(out)))
    return out.value


def mean_recall_at_k(found: np.ndarray, truth_ids: np.ndarray, k: int,
                     counts: np.ndarray | None = None) -> float:
    """Mean recall_at_k over the rows of a [nq, stride] NEIGHBOR_DTYPE results matrix."""
    f = np.ascontiguousarray(found)
    if f.dtype != nat.NEIGHBOR_DTYPE or f.ndim != 2:
        raise ValueError("found must be a [nq, stride] neighbor matrix")
    t = np.ascontiguousarray(truth_ids, dtype=np.int32)
    if t.ndim != 2 or t.shape[0] != f.shape[0]:
        raise ValueError("truth must have one row per query")
    c = None if counts is None else np.ascontiguousarray(counts, dtype=np.uint32)
    out = ctypes.c_double()
    nat.check(nat.lib.vdx_mean_recall_at_k(nat.ptr(f), f.shape[1], nat.ptr(c), nat.ptr(t),
                                           t.shape[1], f.shape[0], int(k), ctypes.byref(out)))
    return out.value


def mean_threshold_recall_at_k(found: np.ndarray, kth_true: np.ndarray, k: int,
                               tolerance: float = 1e-5, counts: np.ndarray | None = None) -> float:
    """Mean threshold_recall_at_k over the rows of a [nq, stride] results matrix."""
    f = np.ascontiguousarray(found)
    if f.dtype != nat.NEIGHBOR_DTYPE or f.ndim != 2:
        raise ValueError("found must be a [nq, stride] neighbor matrix")
    d = np.ascontiguousarray(kth_true, dtype=np.float32)
    if d.shape != (f.shape[0],):
        raise ValueError("kth_true must have one distance per query")
    c = None if counts is None else np.ascontiguousarray(counts, dtype=np.uint32)
    out = ctypes.c_double()
    nat.check(nat.lib.vdx_mean_threshold_recall_at_k(nat.ptr(f), f.shape[1], nat.ptr(c), nat.ptr(d),
                                                     f.shape[0], int(k), float(tolerance),
                                                     ctypes.byref(out)))
    return out.value
