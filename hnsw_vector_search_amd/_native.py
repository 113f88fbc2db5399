"""ctypes binding of libvdxgpu.so (include/vdxgpu.h).

The library is loaded from this package directory only; there is no fallback. If it is missing
the import fails with a message saying how to build it, so nothing ever runs on a CPU
stand-in. VDX_LIBRARY=sanitized selects the same library built with its host code under
ASan + UBSan (`python -m hnsw_vector_search_amd.build --sanitize`), for CPU-side checks run
with the sanitizer runtime preloaded.
"""

from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libvdxgpu.so"
if os.environ.get("VDX_LIBRARY") == "sanitized":
    LIB_PATH = Path(__file__).resolve().parent / "_build_san" / "libvdxgpu.so"

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is not built; run `python -m hnsw_vector_search_amd.build` first")

lib = ctypes.CDLL(str(LIB_PATH))

c_u32, c_u64, c_int, c_float, c_double = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                          ctypes.c_float, ctypes.c_double)
c_ptr, c_char_p = ctypes.c_void_p, ctypes.c_char_p

VDX_OK, VDX_INVALID_ARGUMENT, VDX_RUNTIME_ERROR = 0, 1, 2
VDX_LENGTH_ERROR, VDX_OUT_OF_MEMORY, VDX_DEVICE_ERROR = 3, 4, 5

PRECISIONS = {"fp32": 0, "bf16": 1, "fp8": 2, "exact": 3}
PREC_FP64_TIES, PREC_FP64_ERROR = 0x100, 0x200
STAT_NAMES = ("expansions", "distances", "upper_blocks", "duplicates", "early_stops", "link_prefetched",
              "clamped")
MLP_PARAMS = 5345
NOALLOC_MAX_QUERIES = 1 << 20


class GraphView(ctypes.Structure):
    _fields_ = [("metric", c_u32), ("reserved", c_u32), ("dim", c_u64), ("n", c_u64),
                ("M", c_u64), ("max_level", c_u64), ("entry_point", c_u64),
                ("upper_blocks", c_u64), ("levels", c_ptr), ("layer0", c_ptr), ("upper", c_ptr),
                ("vectors", c_ptr)]


SYNC_FN = ctypes.CFUNCTYPE(c_int, c_ptr, ctypes.POINTER(GraphView), ctypes.POINTER(c_u32), c_u64)
SEARCH_FN = ctypes.CFUNCTYPE(c_int, c_ptr, ctypes.POINTER(c_u32), ctypes.POINTER(c_u32), c_u64, c_u32, c_u32,
                             c_u32, c_ptr, ctypes.POINTER(c_u32))
SEARCHER_SELECTS, SEARCHER_UPPER_LAYERS = 1, 2


class BatchSearcher(ctypes.Structure):
    _fields_ = [("ctx", c_ptr), ("sync", SYNC_FN), ("search", SEARCH_FN), ("caps", c_u32)]


class BuildPlan(ctypes.Structure):
    """vdx_build_plan: host_prefix, batch_min, batch_max, batch_fraction."""
    _fields_ = [("host_prefix", c_u64), ("batch_min", c_u64), ("batch_max", c_u64),
                ("batch_fraction", c_double)]


class BuildStats(ctypes.Structure):
    _fields_ = [("batches", c_u64), ("searched", c_u64), ("host_inserted", c_u64),
                ("host_insert_s", c_double), ("sync_s", c_double), ("search_s", c_double),
                ("link_s", c_double), ("total_s", c_double)]


class GraphReport(ctypes.Structure):
    _fields_ = [("reachable", c_u64), ("max_level", c_u64), ("violation_count", c_u64),
                ("mean_degree0", c_double)]


class SearchOpts(ctypes.Structure):
    _fields_ = [("visited_log2", c_u32), ("waves_per_group", c_u32), ("termination", c_u32),
                ("min_steps", c_u32), ("mlp", c_ptr), ("mlp_threshold", c_float),
                ("reserved_f", c_float), ("stats", c_ptr), ("row_format", c_u32),
                ("mlp_every", c_u32), ("max_expansions", c_u32), ("max_expansions_per_query", c_ptr),
                ("visited_form", c_u32)]


class BatchClass(ctypes.Structure):
    _fields_ = [("ef_cap", c_u32), ("queries", c_u32), ("visited_log2", c_u32),
                ("waves_per_group", c_u32), ("blocks_per_cu", c_u32), ("ms", c_float), ("visited_form", c_u32)]


class BfStats(ctypes.Structure):
    _fields_ = [("shortlist", c_u64), ("widened", c_u64), ("exact_fallback", c_u64),
                ("fp64_ties", c_u64), ("max_abs_error", c_double), ("fp64_reordered", c_u64),
                ("fp64_unresolved", c_u64), ("fp64_window", c_u64), ("max_err_fp64", c_double),
                ("max_rel_err_fp64", c_double)]


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
    "vdx_host_graph_changes": (c_int, [c_ptr, c_u64, c_ptr, c_u64, c_ptr]),
    "vdx_host_graph_save_v2": (c_int, [c_ptr, c_char_p]),
    "vdx_host_graph_add_batched": (c_int, [c_ptr, c_ptr, c_u64, c_u64, c_u64, ctypes.POINTER(BatchSearcher),
                                           ctypes.POINTER(BuildPlan), c_u64, ctypes.POINTER(BuildStats)]),
    "vdx_host_graph_memory_bytes": (c_u64, [c_ptr]),
    "vdx_host_graph_inspect": (c_int, [c_ptr, ctypes.POINTER(GraphReport), c_char_p, c_u64]),
    "vdx_host_graph_free": (None, [c_ptr]),
    "vdx_host_check_new_vectors": (c_int, [c_char_p, c_u32, c_u64, c_u64, c_ptr, c_u64, c_u64]),
    "vdx_host_normalize_rows": (c_int, [c_ptr, c_u64, c_u64]),
    "vdx_host_distance": (c_float, [c_u32, c_ptr, c_ptr, c_u64]),
    "vdx_vecs_info": (c_int, [c_char_p, ctypes.POINTER(c_u64), ctypes.POINTER(c_u64)]),
    "vdx_vecs_read": (c_int, [c_char_p, c_ptr, c_u64, c_u64]),
    "vdx_recall_at_k": (c_int, [c_ptr, c_u64, c_ptr, c_u64, c_u64, ctypes.POINTER(ctypes.c_double)]),
    "vdx_threshold_recall_at_k": (c_int, [c_ptr, c_u64, c_float, c_u64, c_float,
                                          ctypes.POINTER(ctypes.c_double)]),
    "vdx_mean_recall_at_k": (c_int, [c_ptr, c_u64, c_ptr, c_ptr, c_u64, c_u64, c_u64,
                                     ctypes.POINTER(ctypes.c_double)]),
    "vdx_mean_threshold_recall_at_k": (c_int, [c_ptr, c_u64, c_ptr, c_ptr, c_u64, c_u64, c_float,
                                               ctypes.POINTER(ctypes.c_double)]),
    "vdx_gpu_open": (c_int, [c_int, c_ptr]),
    "vdx_gpu_close": (None, [c_ptr]),
    "vdx_gpu_stream": (c_ptr, [c_ptr]),
    "vdx_gpu_alloc": (c_int, [c_ptr, c_u64, c_ptr]),
    "vdx_gpu_free": (None, [c_ptr, c_ptr]),
    "vdx_gpu_copy_to_device": (c_int, [c_ptr, c_ptr, c_ptr, c_u64]),
    "vdx_gpu_copy_to_host": (c_int, [c_ptr, c_ptr, c_ptr, c_u64]),
    "vdx_gpu_copy_device": (c_int, [c_ptr, c_ptr, c_ptr, c_u64]),
    "vdx_gpu_gather_probe": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_u32, c_u32, c_ptr, c_ptr]),
    "vdx_gpu_synchronize": (c_int, [c_ptr]),
    "vdx_gpu_event_record": (c_int, [c_ptr, c_u32]),
    "vdx_gpu_event_elapsed": (c_int, [c_ptr, c_u32, c_u32, ctypes.POINTER(c_float)]),
    "vdx_gpu_hnsw_upload": (c_int, [c_ptr, ctypes.POINTER(GraphView), c_ptr]),
    "vdx_gpu_hnsw_update": (c_int, [c_ptr, ctypes.POINTER(GraphView), c_ptr, c_u64]),
    "vdx_gpu_hnsw_load": (c_int, [c_ptr, c_char_p, c_ptr]),
    "vdx_gpu_hnsw_info": (c_int, [c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]),
    "vdx_gpu_hnsw_u8_rows": (c_int, [c_ptr, c_ptr]),
    "vdx_gpu_hnsw_build_add": (c_int, [c_ptr, c_ptr, c_ptr, c_u64, c_u64, c_u64, ctypes.POINTER(BuildPlan),
                                       c_u64, ctypes.POINTER(BuildStats)]),
    "vdx_gpu_hnsw_free": (None, [c_ptr]),
    "vdx_gpu_hnsw_search": (c_int, [c_ptr, c_ptr, c_u64, c_u64, c_ptr, c_u32, c_ptr, c_u32,
                                    c_ptr, c_u64, c_ptr, c_char_p, ctypes.POINTER(SearchOpts)]),
    "vdx_gpu_hnsw_search_device": (c_int, [c_ptr, c_ptr, c_u64, c_ptr, c_u32, c_ptr, c_u32,
                                           c_u32, c_ptr, c_u64, c_ptr,
                                           ctypes.POINTER(SearchOpts), c_ptr]),
    "vdx_gpu_mlp_create": (c_int, [c_ptr, c_ptr, c_u64, c_ptr]),
    "vdx_gpu_mlp_params": (c_int, [c_ptr, c_ptr, c_u64]),
    "vdx_gpu_mlp_free": (None, [c_ptr]),
    "vdx_gpu_mlp_train": (c_int, [c_ptr, c_ptr, c_ptr, c_u64, c_u32, c_u32, c_float, c_u64, c_ptr]),
    "vdx_gpu_mlp_infer": (c_int, [c_ptr, c_ptr, c_u64, c_ptr]),
    "vdx_gpu_hnsw_trace_device": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_u32, c_ptr, c_u32, c_ptr,
                                          c_ptr, c_ptr, c_ptr]),
    "vdx_gpu_hnsw_batch_create": (c_int, [c_ptr, c_u64, c_u32, c_u32, ctypes.POINTER(SearchOpts), c_ptr]),
    "vdx_gpu_hnsw_batch_free": (None, [c_ptr]),
    "vdx_gpu_hnsw_batch_tune": (c_int, [c_ptr, c_ptr, c_u64, c_ptr, c_u32, c_ptr, c_u32, c_ptr, c_u64,
                                        c_ptr, c_u32]),
    "vdx_gpu_hnsw_batch_run": (c_int, [c_ptr, c_ptr, c_u64, c_ptr, c_u32, c_ptr, c_u32, c_ptr, c_u64,
                                       c_ptr, c_int, c_ptr]),
    "vdx_gpu_hnsw_batch_classes": (c_int, [c_ptr, ctypes.POINTER(BatchClass), c_u32,
                                           ctypes.POINTER(c_u32)]),
    "vdx_gpu_hnsw_search_uneven_noalloc": (c_int, [c_ptr, c_ptr, c_u64, c_ptr, c_u32, c_ptr, c_u32, c_u32, c_u32,
                                                   c_ptr, c_u64, c_ptr, ctypes.POINTER(SearchOpts), c_ptr]),
    "vdx_gpu_bf_upload": (c_int, [c_ptr, c_u32, c_u64, c_u64, c_ptr, c_ptr]),
    "vdx_gpu_bf_append": (c_int, [c_ptr, c_ptr, c_u64, c_u64]),
    "vdx_gpu_bf_free": (None, [c_ptr]),
    "vdx_gpu_bf_search": (c_int, [c_ptr, c_ptr, c_u64, c_u64, c_u32, c_u32, c_ptr, c_u64,
                                  c_ptr, c_char_p, ctypes.POINTER(BfStats)]),
    "vdx_gpu_bf_gemm_shape": (c_int, [c_ptr, c_u32]),
    "vdx_gpu_bf_tune": (c_int, [c_ptr, c_ptr, c_u64, c_u32, c_u32, c_ptr, c_ptr]),
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


class DeviceArray:
    """An HBM buffer owned through the library (one HIP runtime per process), shaped like a
    numpy array for copies in and out."""

    def __init__(self, shape, dtype, ctx: "Context | None" = None):
        self.ctx = ctx or Context.get(0)
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = c_ptr()
        check(lib.vdx_gpu_alloc(self.ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray, ctx: "Context | None" = None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, ctx)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib.vdx_gpu_copy_to_device(self.ctx.handle, self.ptr, a.ctypes.data, self.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib.vdx_gpu_copy_to_host(self.ctx.handle, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            lib.vdx_gpu_free(self.ctx.handle, self.ptr)
            self.ptr = None


def elapsed_ms(ctx: Context, start_slot: int, stop_slot: int) -> float:
    ms = c_float()
    check(lib.vdx_gpu_event_elapsed(ctx.handle, start_slot, stop_slot, ctypes.byref(ms)))
    return ms.value
