"""TEXMEX .fvecs / .ivecs files (vecdex src/io.cpp:30-87): readers through the library's C++
reader (vecdex's checks and messages; RuntimeError on any I/O or format error), and writers for
exporting synthetic workloads in the same format."""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _native as nat


def _read(path, dtype) -> np.ndarray:
    p = os.fsencode(os.fspath(path))
    rows, dim = nat.c_u64(), nat.c_u64()
    nat.check(nat.lib.vdx_vecs_info(p, ctypes.byref(rows), ctypes.byref(dim)))
    out = np.empty((rows.value, dim.value), dtype=dtype)
    nat.check(nat.lib.vdx_vecs_read(p, nat.ptr(out), rows.value, dim.value))
    return out


def read_fvecs(path) -> np.ndarray:
    """float32 [rows, dim] (read_fvecs, include/vecdex/io.hpp:18)."""
    return _read(path, np.float32)


def read_ivecs(path) -> np.ndarray:
    """int32 [rows, dim] (read_ivecs, include/vecdex/io.hpp:19)."""
    return _read(path, np.int32)


def _write(path, rows: np.ndarray, dtype) -> None:
    a = np.ascontiguousarray(rows, dtype=dtype)
    if a.ndim != 2 or a.shape[1] == 0:
        raise ValueError("need a non-empty [rows, dim] matrix")
    rec = np.empty((a.shape[0], a.shape[1] + 1), dtype=np.int32)
    rec[:, 0] = a.shape[1]
    rec[:, 1:] = a.view(np.int32)
    rec.tofile(os.fspath(path))


def write_fvecs(path, rows) -> None:
    _write(path, rows, np.float32)


def write_ivecs(path, rows) -> None:
    _write(path, rows, np.int32)
