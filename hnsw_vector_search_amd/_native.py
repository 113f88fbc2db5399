"""ctypes binding of libvdxgpu.so (include/vdxgpu.h).

The library is loaded from this package directory only; there is no fallback. If it is missing
the import fails with a message saying how to build it, so nothing ever runs on a CPU
stand-in.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libvdxgpu.so"

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
        "or `python -m hnsw_vector_search_amd.build` first")

lib = ctypes.CDLL(str(LIB_PATH))

c_u32, c_u64, c_int, c_float, c_double = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                          ctypes.c_float, ctypes.c_double)
c_ptr, c_char_p = ctypes.c_void_p, ctypes.c_char_p

VDX_OK, VDX_INVALID_ARGUMENT, VDX_RUNTIME_ERROR = 0, 1, 2
VDX_LENGTH_ERROR, VDX_OUT_OF_MEMORY, VDX_DEVICE_ERROR = 3, 4, 5

PRECISIONS = {"fp32": 0, "bf16": 1, "fp8": 2, "exact": 3}
STAT_NAMES = ("expansions", "distances", "upper_blocks", "duplicates")


class GraphView(ctypes.Structure):
    _fields_ = [("metric", c_u32), ("reserved", c_u32), ("dim", c_u64), ("n", c_u64),
                ("M", c_u64), ("max_level", c_u64), ("entry_point", c_u64),
                ("upper_blocks", c_u64), ("levels", c_ptr), ("layer0", c_ptr), ("upper", c_ptr),
                ("vectors", c_ptr)]


class GraphReport(ctypes.Structure):
    _fields_ = [("reachable", c_u64), ("max_level", c_u64), ("violation_count", c_u64),
                ("mean_degree0", c_double)]


class SearchOpts(ctypes.Structure):
    _fields_ = [("visited_log2", c_u32), ("waves_per_group", c_u32), ("termination", c_u32),
                ("reserved", c_u32), ("mlp", c_ptr), ("mlp_threshold", c_float),
                ("reserved_f", c_float), ("stats", c_ptr)]


class BfStats(ctypes.Structure):
    _fields_ = [("shortlist", c_u64), ("widened", c_u64), ("exact_fallback", c_u64),
                ("fp64_ties", c_u64), ("max_abs_error", c_double)]


NEIGHBOR_DTYPE = np.dtype([("distance", np.float32), ("id", np.uint32)])

_SIGS = {
    "vdx_last_error": (c_char_p, []),
    "vdx_version": (c_char_p, []),
    "vdx_host_graph_create": (c_int, [c_u64, c_u32, c_u64, c_u64, c_u64, c_ptr]),
    "vdx_host_graph_add": (c_int, [c_ptr, c_ptr, c_u64, c_u64, c_u64]),
    "vdx_host_graph_save": (c_int, [c_ptr, c_char_p]),
    "vdx_host_graph_load": (c_int, [c_char_p, c_ptr]),
    "vdx_host_graph_view": (c_int, [c_ptr, ctypes.POINTER(GraphView)]),
    "vdx_host_graph_params": (c_int, [c_ptr, c_ptr, c_ptr, c_ptr]),
    "vdx_host_graph_memory_bytes": (c_u64, [c_ptr]),
    "vdx_host_graph_inspect": (c_int, [c_ptr, ctypes.POINTER(GraphReport), c_char_p, c_u64]),
    "vdx_host_graph_free": (None, [c_ptr]),
    "vdx_host_check_new_vectors": (c_int, [c_char_p, c_u32, c_u64, c_u64, c_ptr, c_u64, c_u64]),
    "vdx_host_normalize_rows": (c_int, [c_ptr, c_u64, c_u64]),
    "vdx_host_distance": (c_float, [c_u32, c_ptr, c_ptr, c_u64]),
    "vdx_gpu_open": (c_int, [c_int, c_ptr]),
    "vdx_gpu_close": (None, [c_ptr]),
    "vdx_gpu_stream": (c_ptr, [c_ptr]),
    "vdx_gpu_hnsw_upload": (c_int, [c_ptr, ctypes.POINTER(GraphView), c_ptr]),
    "vdx_gpu_hnsw_free": (None, [c_ptr]),
    "vdx_gpu_hnsw_search": (c_int, [c_ptr, c_ptr, c_u64, c_u64, c_ptr, c_u32, c_ptr, c_u32,
                                    c_ptr, c_u64, c_ptr, c_char_p, ctypes.POINTER(SearchOpts)]),
    "vdx_gpu_hnsw_search_device": (c_int, [c_ptr, c_ptr, c_u64, c_ptr, c_u32, c_ptr, c_u32,
                                           c_u32, c_ptr, c_u64, c_ptr,
                                           ctypes.POINTER(SearchOpts), c_ptr]),
    "vdx_gpu_bf_upload": (c_int, [c_ptr, c_u32, c_u64, c_u64, c_ptr, c_ptr]),
    "vdx_gpu_bf_free": (None, [c_ptr]),
    "vdx_gpu_bf_search": (c_int, [c_ptr, c_ptr, c_u64, c_u64, c_u32, c_u32, c_ptr, c_u64,
                                  c_ptr, c_char_p, ctypes.POINTER(BfStats)]),
    "vdx_gpu_bf_search_device": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_u32, c_ptr, c_u64,
                                         c_ptr, ctypes.POINTER(BfStats)]),
}

for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED = tuple(_SIGS)


def check(status: int) -> None:
    """Raises the Python exception vecdex's binding raises for the C++ exception type."""
    if status == VDX_OK:
        return
    msg = (lib.vdx_last_error() or b"").decode("utf-8", "replace")
    if status in (VDX_INVALID_ARGUMENT, VDX_LENGTH_ERROR):
        raise ValueError(msg)
    if status == VDX_OUT_OF_MEMORY:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


class Context:
    """One HIP device, opened once per process (vdx_gpu_open)."""

    _lock = threading.Lock()
    _contexts: dict[int, "Context"] = {}

    def __init__(self, device: int):
        handle = c_ptr()
        check(lib.vdx_gpu_open(device, ctypes.byref(handle)))
        self.handle = handle
        self.device = device

    @classmethod
    def get(cls, device: int = 0) -> "Context":
        with cls._lock:
            ctx = cls._contexts.get(device)
            if ctx is None:
                ctx = cls._contexts[device] = Context(device)
            return ctx

    @property
    def stream(self) -> int:
        return lib.vdx_gpu_stream(self.handle) or 0
