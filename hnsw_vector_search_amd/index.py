"""The host-side mirror of vecdex's interfaces for the query path, over the C ABI.

* HnswIndex / BruteForceIndex: vecdex's C++ classes (include/vecdex/hnsw.hpp:63-184,
  include/vecdex/brute_force.hpp:21-43) with the same method names, argument meanings, result
  order and exception types. search/search_batch run on the GPU.
* Index: vecdex's Python binding (python/bindings.cpp:190-417, python/vecdex/_core.pyi):
  zero-copy float32 inputs with the same input rules and messages, int64 ids and float32
  distances, -1 / inf padding, ef default 64, a reader-writer lock (searches together, adds
  alone).

The graph is built or loaded on the host (restated vecdex build, FORMAT v1) and copied to HBM
on the first search after any change.
"""

from __future__ import annotations

import ctypes
import os
import threading
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _native as nat

SPACES = {"l2": 0, "ip": 1, "cosine": 2}
SPACE_NAMES = {v: k for k, v in SPACES.items()}
DEFAULT_EF = 64  # python/bindings.cpp:42


class Metric:
    """vecdex::Metric (include/vecdex/distance.hpp:9-14) as FORMAT codes."""
    L2 = 0
    InnerProduct = 1
    Cosine = 2


@dataclass(frozen=True)
class Neighbor:
    """vecdex::Neighbor: distance under the metric, then id."""
    distance: float
    id: int


@dataclass
class GraphReport:
    violations: list = field(default_factory=list)
    reachable: int = 0
    max_level: int = 0
    mean_degree0: float = 0.0


class _RWLock:
    """Many readers or one writer (the binding's std::shared_mutex, python/bindings.cpp:312).
    A waiting writer holds off new readers: searches that re-take the lock back to back would
    otherwise keep the reader count above zero and starve add_items forever."""

    def __init__(self):
        self._cond = threading.Condition()
        self._readers = 0
        self._writer = False
        self._writers_waiting = 0

    def acquire_read(self):
        with self._cond:
            while self._writer or self._writers_waiting:
                self._cond.wait()
            self._readers += 1

    def release_read(self):
        with self._cond:
            self._readers -= 1
            if self._readers == 0:
                self._cond.notify_all()

    def acquire_write(self):
        with self._cond:
            self._writers_waiting += 1
            try:
                while self._writer or self._readers:
                    self._cond.wait()
            finally:
                self._writers_waiting -= 1
            self._writer = True

    def release_write(self):
        with self._cond:
            self._writer = False
            self._cond.notify_all()


def _positive(value, name: str) -> int:
    value = int(value)
    if value < 1:
        raise ValueError(f"{name} must be positive, got {value}")
    return value


def _thread_count(value) -> int:
    value = int(value)
    if value < 0:
        raise ValueError(
            f"num_threads must be positive, or 0 for one per hardware thread, got {value}")
    return value


def _rows(data, dim: int, method: str) -> np.ndarray:
    """A float32 (n, dim) view of `data` without copying it (python/bindings.cpp:139-181)."""
    arr = data if isinstance(data, np.ndarray) else np.asarray(memoryview(data))
    if arr.ndim not in (1, 2):
        raise ValueError(f"{method}: data must be 1-D (one vector) or 2-D (one vector per row), "
                         f"got {arr.ndim}-D")
    if arr.dtype != np.float32 or not arr.dtype.isnative:
        raise TypeError(f"{method}: data must be float32, got {arr.dtype}; "
                        "data.astype(np.float32) converts it (a copy)")
    if arr.shape[-1] != dim:
        raise ValueError(f"{method}: vectors have dimension {arr.shape[-1]}, the index has {dim}")
    if not arr.flags.c_contiguous:
        raise ValueError(f"{method}: data must be C-contiguous; np.ascontiguousarray(data) "
                         "makes a contiguous copy")
    if arr.size and arr.ctypes.data % 4 != 0:
        raise ValueError(f"{method}: data is not aligned for float32; data.copy() makes an "
                         "aligned copy")
    return arr.reshape(-1, dim) if arr.ndim == 1 else arr


def _as_matrix(values, dim: int | None = None) -> np.ndarray:
    arr = np.ascontiguousarray(values, dtype=np.float32)
    if arr.ndim == 1:
        arr = arr.reshape(1, -1) if dim is None else arr.reshape(-1, dim)
    return arr


class _HostGraph:
    """Owns a vdx_host_graph."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def create(cls, dim, metric, M, ef_construction, seed):
        h = ctypes.c_void_p()
        nat.check(nat.lib.vdx_host_graph_create(dim, metric, M, ef_construction, seed,
                                                ctypes.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, path):
        h = ctypes.c_void_p()
        nat.check(nat.lib.vdx_host_graph_load(os.fsencode(os.fspath(path)), ctypes.byref(h)))
        return cls(h)

    def __del__(self):
        if getattr(self, "handle", None):
            nat.lib.vdx_host_graph_free(self.handle)
            self.handle = None

    def view(self) -> nat.GraphView:
        v = nat.GraphView()
        nat.check(nat.lib.vdx_host_graph_view(self.handle, ctypes.byref(v)))
        return v

    def params(self):
        efc, seed, version = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        nat.check(nat.lib.vdx_host_graph_params(self.handle, ctypes.byref(efc), ctypes.byref(seed),
                                                ctypes.byref(version)))
        return efc.value, seed.value, version.value

    def changes(self, since: int) -> np.ndarray:
        """Ascending ids of the nodes whose link blocks the adds after version `since` wrote."""
        count = ctypes.c_uint64()
        nat.check(nat.lib.vdx_host_graph_changes(self.handle, since, None, 0, ctypes.byref(count)))
        out = np.zeros(count.value, dtype=np.uint32)
        nat.check(nat.lib.vdx_host_graph_changes(self.handle, since, nat.ptr(out), out.size,
                                                 ctypes.byref(count)))
        return out


class _GpuSearch:
    """The search methods shared by the host-graph index and the device-only index; the class
    provides gpu_handle()."""

    def search_arrays(self, queries: np.ndarray, k, ef, *, where: str = "HnswIndex::search_batch",
                      opts: nat.SearchOpts | None = None):
        """GPU search of prepared-or-raw host queries; k and ef may be ints or per-query
        arrays. Returns (neighbors[nq, kmax] structured array, counts[nq])."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 1:
            q = q.reshape(1, -1)
        nq = q.shape[0]
        ks = None if np.isscalar(k) else np.ascontiguousarray(k, dtype=np.uint32)
        efs = None if np.isscalar(ef) else np.ascontiguousarray(ef, dtype=np.uint32)
        kmax = int(k) if ks is None else int(ks.max(initial=1))
        # every entry is written (rows past their count zeroed)
        out = np.empty((nq, max(kmax, 1)), dtype=nat.NEIGHBOR_DTYPE)
        count = np.empty(nq, dtype=np.uint32)
        if nq == 0:
            return out, count
        handle = self.gpu_handle()
        nat.check(nat.lib.vdx_gpu_hnsw_search(
            handle, nat.ptr(q), nq, q.shape[1], nat.ptr(ks), 0 if ks is not None else int(k),
            nat.ptr(efs), 0 if efs is not None else int(ef), nat.ptr(out), out.shape[1],
            nat.ptr(count), where.encode(), ctypes.byref(opts) if opts is not None else None))
        return out, count

    def gpu_u8_rows(self) -> bool:
        """Whether the device copy also holds its rows as bytes (dimension 128, every value an
        integer in [0, 255]); the plain search then reads those (same results, a quarter of the
        row traffic) unless opts.row_format = 1 (vdx_gpu_hnsw_u8_rows)."""
        kept = ctypes.c_uint32(0)
        nat.check(nat.lib.vdx_gpu_hnsw_u8_rows(self.gpu_handle(), ctypes.byref(kept)))
        return bool(kept.value)

    def search(self, query: Sequence[float], k: int, ef: int) -> list:
        """Up to k approximate nearest neighbors, closest first (HnswIndex::search)."""
        q = np.asarray(query, dtype=np.float32).reshape(1, -1)
        if k == 0:
            raise ValueError("HnswIndex::search: k must be positive")
        out, count = self.search_arrays(q, k, ef, where="HnswIndex::search")
        return [Neighbor(float(d), int(i)) for d, i in out[0][: count[0]]]

    def search_batch(self, queries, k: int, ef: int, num_threads: int = 0) -> list:
        """search() for every row (HnswIndex::search_batch); num_threads is accepted and has no
        effect on results."""
        q = _as_matrix(queries)
        if k == 0 and q.shape[0]:
            raise ValueError("HnswIndex::search_batch: k must be positive")
        out, count = self.search_arrays(q, k, ef, where="HnswIndex::search_batch")
        return [[Neighbor(float(d), int(i)) for d, i in out[r][: count[r]]]
                for r in range(q.shape[0])]


class HnswIndex(_GpuSearch):
    """vecdex::HnswIndex with search/search_batch on the GPU.

    HnswIndex(dim, metric=Metric.L2, M=16, ef_construction=200, seed=42)
    (include/vecdex/hnsw.hpp:28-41, 67).
    """

    def __init__(self, dim: int, metric: int = Metric.L2, M: int = 16,
                 ef_construction: int = 200, seed: int = 42, *, device: int = 0,
                 _graph: _HostGraph | None = None):
        self._graph = _graph or _HostGraph.create(dim, metric, M, ef_construction, seed)
        self._device = device
        self._gpu = None
        self._gpu_version = None
        self._gpu_lock = threading.Lock()
        self.last_build_stats = None  # vdx_build_stats of the last gpu_assisted add

    # ---- properties -------------------------------------------------------------------
    def _view(self) -> nat.GraphView:
        return self._graph.view()

    def dim(self) -> int:
        return self._view().dim

    def size(self) -> int:
        return self._view().n

    def metric(self) -> int:
        return self._view().metric

    def params(self) -> dict:
        v = self._view()
        efc, seed, _ = self._graph.params()
        return {"M": v.M, "ef_construction": efc, "seed": seed}

    def __len__(self) -> int:
        return self.size()

    # ---- build (host) -----------------------------------------------------------------
    def add(self, vectors, num_threads: int = 1, *, gpu_assisted: bool = False,
            plan: nat.BuildPlan | None = None) -> None:
        """Stores and links every row (HnswIndex::add, src/hnsw.cpp:259-291). A 1-D input is
        one vector. gpu_assisted: the layer-0 insertion searches of level-0 nodes run on the GPU
        in batches (vdx_gpu_hnsw_build_add, batch sizes from `plan`); nodes of one batch are not
        candidates of each other, so the graph differs from the host build's (a batch of one
        node on one thread builds the same graph)."""
        arr = _as_matrix(vectors)
        if gpu_assisted:
            handle = self.gpu_handle()
            stats = nat.BuildStats()
            with self._gpu_lock:
                nat.check(nat.lib.vdx_gpu_hnsw_build_add(
                    handle, self._graph.handle, nat.ptr(arr), arr.shape[0], arr.shape[1], num_threads,
                    ctypes.byref(plan) if plan is not None else None, self._gpu_version,
                    ctypes.byref(stats)))
                self._gpu_version = self._graph.params()[2]
            self.last_build_stats = {name: getattr(stats, name) for name, _ in stats._fields_}
            return
        nat.check(nat.lib.vdx_host_graph_add(self._graph.handle, nat.ptr(arr), arr.shape[0],
                                             arr.shape[1], num_threads))
        self._refresh_gpu()

    def _refresh_gpu(self) -> None:
        """Brings an existing HBM copy up to the host graph at once (an index never searched on
        the GPU uploads on its first search)."""
        if self._gpu is not None:
            self.gpu_handle()

    def inspect(self) -> GraphReport:
        r = nat.GraphReport()
        buf = ctypes.create_string_buffer(512)
        nat.check(nat.lib.vdx_host_graph_inspect(self._graph.handle, ctypes.byref(r), buf, 512))
        out = GraphReport(reachable=r.reachable, max_level=r.max_level,
                          mean_degree0=r.mean_degree0)
        if r.violation_count:
            out.violations = [buf.value.decode()] + [""] * (r.violation_count - 1)
        return out

    def memory_bytes(self) -> int:
        return nat.lib.vdx_host_graph_memory_bytes(self._graph.handle)

    def save(self, path, version: int = 1) -> None:
        """HnswIndex::save (FORMAT v1), or version=2: this library's aligned, mappable format,
        which HnswDeviceIndex.load streams straight to HBM. load() reads both."""
        if version not in (1, 2):
            raise ValueError(f"HnswIndex::save: format version {version} is not 1 or 2")
        fn = nat.lib.vdx_host_graph_save if version == 1 else nat.lib.vdx_host_graph_save_v2
        nat.check(fn(self._graph.handle, os.fsencode(os.fspath(path))))

    @classmethod
    def load(cls, path, *, device: int = 0) -> "HnswIndex":
        return cls(0, _graph=_HostGraph.load(path), device=device)

    def graph_arrays(self) -> dict:
        """Zero-copy numpy views of the FORMAT v1 sections (valid until the next add)."""
        v = self._view()
        def arr(p, count, dtype):
            if count == 0 or not p:
                return np.zeros(0, dtype=dtype)
            buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(p)
            return np.frombuffer(buf, dtype=dtype, count=count)
        return {
            "metric": v.metric, "dim": v.dim, "n": v.n, "M": v.M, "max_level": v.max_level,
            "entry_point": v.entry_point, "upper_blocks": v.upper_blocks,
            "levels": arr(v.levels, v.n, np.uint8),
            "layer0": arr(v.layer0, v.n * (1 + 2 * v.M), np.uint32),
            "upper": arr(v.upper, v.upper_blocks * (1 + v.M), np.uint32),
            "vectors": arr(v.vectors, v.n * v.dim, np.float32).reshape(v.n, v.dim),
        }

    # ---- GPU copy ---------------------------------------------------------------------
    def gpu_handle(self):
        """The HBM copy of the graph: uploaded on first use, then brought up to date after each
        add by appending the new nodes and rewriting the blocks the add linked back into
        (vdx_gpu_hnsw_update), so the handle stays the same for the index's lifetime."""
        _, _, version = self._graph.params()
        with self._gpu_lock:
            if self._gpu is None:
                ctx = nat.Context.get(self._device)
                h = ctypes.c_void_p()
                view = self._view()
                nat.check(nat.lib.vdx_gpu_hnsw_upload(ctx.handle, ctypes.byref(view),
                                                      ctypes.byref(h)))
                self._gpu, self._gpu_version = h, version
            elif self._gpu_version != version:
                changed = self._graph.changes(self._gpu_version)
                view = self._view()
                nat.check(nat.lib.vdx_gpu_hnsw_update(self._gpu, ctypes.byref(view),
                                                      nat.ptr(changed), changed.size))
                self._gpu_version = version
            return self._gpu

    def __del__(self):
        if getattr(self, "_gpu", None):
            nat.lib.vdx_gpu_hnsw_free(self._gpu)
            self._gpu = None

class HnswDeviceIndex(_GpuSearch):
    """A graph held in HBM only, loaded from a FORMAT v2 file (vdx_gpu_hnsw_load): the file is
    mapped and each section streamed to the device through pinned chunks, every chunk checked
    before its copy; no host copy of the graph is made. Serves the same searches as HnswIndex
    (search, search_batch, search_arrays, SearchBatch); adding needs the host graph
    (HnswIndex.load)."""

    def __init__(self, handle, device: int):
        self._gpu = handle
        self._device = device

    @classmethod
    def load(cls, path, device: int = 0) -> "HnswDeviceIndex":
        ctx = nat.Context.get(device)
        h = ctypes.c_void_p()
        nat.check(nat.lib.vdx_gpu_hnsw_load(ctx.handle, os.fsencode(os.fspath(path)), ctypes.byref(h)))
        return cls(h, device)

    def gpu_handle(self):
        return self._gpu

    def _info(self):
        metric, dim, n, M = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        nat.check(nat.lib.vdx_gpu_hnsw_info(self._gpu, ctypes.byref(metric), ctypes.byref(dim), ctypes.byref(n),
                                            ctypes.byref(M)))
        return metric.value, dim.value, n.value, M.value

    def dim(self) -> int:
        return self._info()[1]

    def size(self) -> int:
        return self._info()[2]

    def metric(self) -> int:
        return self._info()[0]

    def __len__(self) -> int:
        return self.size()

    def __del__(self):
        if getattr(self, "_gpu", None):
            nat.lib.vdx_gpu_hnsw_free(self._gpu)
            self._gpu = None


class SearchBatch:
    """Uneven batches on the GPU: every query with its own (ef, k), served by one persistent
    launch per R-capacity class, optionally replayed as a captured HIP graph
    (vdx_gpu_hnsw_batch_*). Buffers are device arrays (nat.DeviceArray); queries must be
    prepared (unit length under cosine). Each query's result equals HnswIndex.search at its
    (ef, k). Keep the index alive while the batch is in use."""

    def __init__(self, index: "HnswIndex | HnswDeviceIndex", max_nq: int, max_k: int, max_ef: int,
                 opts: nat.SearchOpts | None = None):
        self.index = index
        self.max_k = int(max_k)
        self._opts = opts  # a learned model referenced by opts must stay alive
        h = ctypes.c_void_p()
        nat.check(nat.lib.vdx_gpu_hnsw_batch_create(index.gpu_handle(), int(max_nq), int(max_k),
                                                    int(max_ef),
                                                    ctypes.byref(opts) if opts is not None else None,
                                                    ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            nat.lib.vdx_gpu_hnsw_batch_free(self._h)
            self._h = None

    @staticmethod
    def _args(dq, nq, ks, k, efs, ef, out, count):
        return (dq.ptr, int(nq), None if ks is None else ks.ptr, int(k),
                None if efs is None else efs.ptr, int(ef), out.ptr, out.shape[1], count.ptr)

    def tune(self, dq, nq, out, count, *, ks=None, k=0, efs=None, ef=0, reps: int = 3) -> list:
        """Times each class's launch shapes on this sample and keeps the fastest."""
        nat.check(nat.lib.vdx_gpu_hnsw_batch_tune(self._h, *self._args(dq, nq, ks, k, efs, ef, out, count),
                                                  int(reps)))
        return self.classes()

    def run(self, dq, nq, out, count, *, ks=None, k=0, efs=None, ef=0, graph: bool = True,
            stream=None) -> None:
        """Enqueues the batch (asynchronous)."""
        nat.check(nat.lib.vdx_gpu_hnsw_batch_run(self._h, *self._args(dq, nq, ks, k, efs, ef, out, count),
                                                 1 if graph else 0, stream))

    def classes(self) -> list:
        buf = (nat.BatchClass * 8)()
        n = ctypes.c_uint32()
        nat.check(nat.lib.vdx_gpu_hnsw_batch_classes(self._h, buf, 8, ctypes.byref(n)))
        return [{name: getattr(buf[i], name) for name, _ in nat.BatchClass._fields_}
                for i in range(n.value)]


class BruteForceIndex:
    """vecdex::BruteForceIndex (include/vecdex/brute_force.hpp:21-43) with a batched GPU search.

    search(query, k) returns the exact top-k (bit-equal to vecdex's). `precision` picks the
    GPU method: an MFMA shortlist ("bf16", the default, "fp32", "fp8") re-ranked exactly and
    certified, or "exact" (every distance in the exact order); all return the same result.
    """

    def __init__(self, dim: int, metric: int = Metric.L2, *, device: int = 0):
        if dim == 0:
            raise ValueError("Matrix: dim must be positive")
        self._dim = int(dim)
        self._metric = int(metric)
        self._device = device
        self._vectors = np.zeros((0, self._dim), dtype=np.float32)
        self._gpu = None
        self._gpu_rows = -1
        self.last_stats = nat.BfStats()

    def dim(self) -> int:
        return self._dim

    def size(self) -> int:
        return self._vectors.shape[0]

    def metric(self) -> int:
        return self._metric

    def __len__(self) -> int:
        return self.size()

    def add(self, vectors) -> int | None:
        """Stores rows (normalized under cosine); a 1-D input is one vector and its id is
        returned (src/brute_force.cpp:43-66)."""
        one = np.ndim(vectors) == 1
        arr = np.array(vectors, dtype=np.float32, ndmin=2, copy=True)
        if one:
            arr = arr.reshape(1, -1)
        nat.check(nat.lib.vdx_host_check_new_vectors(
            b"BruteForceIndex::add", self._metric, self._dim, self.size(), nat.ptr(arr),
            arr.shape[0], arr.shape[1]))
        if self._metric == Metric.Cosine:
            nat.check(nat.lib.vdx_host_normalize_rows(nat.ptr(arr), arr.shape[0], arr.shape[1]))
        first = self.size()
        self._vectors = np.concatenate([self._vectors, arr]) if first else arr
        return first if one else None

    def vectors(self) -> np.ndarray:
        return self._vectors

    def gpu_handle(self):
        """The HBM copy of the stored rows: uploaded on first use; rows added since are appended
        (vdx_gpu_bf_append)."""
        if self._gpu is None:
            ctx = nat.Context.get(self._device)
            h = ctypes.c_void_p()
            v = np.ascontiguousarray(self._vectors)
            nat.check(nat.lib.vdx_gpu_bf_upload(ctx.handle, self._metric, self._dim, v.shape[0],
                                                nat.ptr(v), ctypes.byref(h)))
            self._gpu, self._gpu_rows = h, self.size()
        elif self._gpu_rows != self.size():
            new = np.ascontiguousarray(self._vectors[self._gpu_rows:])
            nat.check(nat.lib.vdx_gpu_bf_append(self._gpu, nat.ptr(new), new.shape[0], self._dim))
            self._gpu_rows = self.size()
        return self._gpu

    def __del__(self):
        if getattr(self, "_gpu", None):
            nat.lib.vdx_gpu_bf_free(self._gpu)
            self._gpu = None

    def set_gemm_shape(self, shape: int) -> None:
        """The BF16 filter GEMM's shape for 97-128-dimensional rows (vdx_gpu_bf_gemm_shape): 0,
        one 32-query block per wave; 1, two. Results are the same either way."""
        nat.check(nat.lib.vdx_gpu_bf_gemm_shape(self.gpu_handle(), int(shape)))

    def tune(self, queries, k: int = 10, reps: int = 3) -> dict:
        """Times the BF16 search in both GEMM shapes on these queries (prepared: normalized under
        cosine) and keeps the faster (vdx_gpu_bf_tune)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        dq = nat.DeviceArray.from_numpy(q, nat.Context.get(self._device))
        ms = (ctypes.c_float * 2)()
        shape = ctypes.c_uint32(0)
        nat.check(nat.lib.vdx_gpu_bf_tune(self.gpu_handle(), dq.ptr, q.shape[0], int(k), int(reps), ms,
                                          ctypes.byref(shape)))
        return {"ms": [float(ms[0]), float(ms[1])], "shape": int(shape.value)}

    def search_arrays(self, queries, k: int, precision: str = "bf16",
                      where: str = "BruteForceIndex::search", fp64: str | None = None):
        """fp64: None; "error" also measures the returned distances against float64 on the GPU
        (last_stats.max_err_fp64); "ties" re-ranks candidates whose float32 distances chain
        within one ulp by their float64 distances (and measures)."""
        flags = {None: 0, "error": nat.PREC_FP64_ERROR, "ties": nat.PREC_FP64_TIES}[fp64]
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 1:
            q = q.reshape(1, -1)
        nq = q.shape[0]
        keep = max(1, min(int(k), self.size()))
        out = np.zeros((nq, keep), dtype=nat.NEIGHBOR_DTYPE)
        count = np.zeros(nq, dtype=np.uint32)
        if nq == 0:
            return out, count
        self.last_stats = nat.BfStats()
        nat.check(nat.lib.vdx_gpu_bf_search(
            self.gpu_handle(), nat.ptr(q), nq, q.shape[1], int(k), nat.PRECISIONS[precision] | flags,
            nat.ptr(out), keep, nat.ptr(count), where.encode(), ctypes.byref(self.last_stats)))
        return out, count

    def search(self, query, k: int, precision: str = "bf16") -> list:
        out, count = self.search_arrays(np.asarray(query, dtype=np.float32).reshape(1, -1), k,
                                        precision)
        return [Neighbor(float(d), int(i)) for d, i in out[0][: count[0]]]

    def search_batch(self, queries, k: int, precision: str = "bf16") -> list:
        q = _as_matrix(queries, self._dim)
        out, count = self.search_arrays(q, k, precision)
        return [[Neighbor(float(d), int(i)) for d, i in out[r][: count[r]]]
                for r in range(q.shape[0])]


class Index:
    """vecdex.Index (python/bindings.cpp:317-417): knn_query runs on the GPU."""

    def __init__(self, space: str, dim: int, M: int = 16, ef_construction: int = 200,
                 seed: int = 42, *, device: int = 0, _core: HnswIndex | None = None):
        if _core is None:
            if space not in SPACES:
                raise ValueError(f"space must be 'l2', 'ip' or 'cosine', got '{space}'")
            _core = HnswIndex(_positive(dim, "dim"), SPACES[space], _positive(M, "M"),
                              _positive(ef_construction, "ef_construction"), int(seed),
                              device=device)
        self._core = _core
        self._ef = DEFAULT_EF
        self._lock = _RWLock()

    # fixed at construction
    @property
    def dim(self) -> int:
        return self._core.dim()

    @property
    def space(self) -> str:
        return SPACE_NAMES[self._core.metric()]

    @property
    def M(self) -> int:
        return self._core.params()["M"]

    @property
    def ef_construction(self) -> int:
        return self._core.params()["ef_construction"]

    @property
    def ef(self) -> int:
        return self._ef

    @ef.setter
    def ef(self, value) -> None:
        self._ef = _positive(value, "ef")

    def set_ef(self, ef) -> None:
        self.ef = ef

    def __len__(self) -> int:
        self._lock.acquire_read()
        try:
            return self._core.size()
        finally:
            self._lock.release_read()

    def memory_bytes(self) -> int:
        self._lock.acquire_read()
        try:
            return self._core.memory_bytes()
        finally:
            self._lock.release_read()

    def add_items(self, data, *, num_threads: int = 1) -> np.ndarray:
        rows = _rows(data, self.dim, "add_items")
        threads = _thread_count(num_threads)
        ids = np.empty(rows.shape[0], dtype=np.int64)
        self._lock.acquire_write()
        try:
            first = self._core.size()
            if rows.shape[0]:
                nat.check(nat.lib.vdx_host_graph_add(self._core._graph.handle, nat.ptr(rows),
                                                     rows.shape[0], rows.shape[1], threads))
                # the device copy follows under the same write lock, so no search sees it stale
                self._core._refresh_gpu()
        finally:
            self._lock.release_write()
        ids[:] = np.arange(first, first + rows.shape[0], dtype=np.int64)
        return ids

    def knn_query(self, data, k: int = 1, ef: int | None = None, *, num_threads: int = 1):
        queries = _rows(data, self.dim, "knn_query")
        want = _positive(k, "k")
        breadth = _positive(ef, "ef") if ef is not None else self._ef
        threads = _thread_count(num_threads)
        n = queries.shape[0]
        ids = np.full((n, want), -1, dtype=np.int64)
        distances = np.full((n, want), np.inf, dtype=np.float32)
        if n == 0:
            return ids, distances
        where = "HnswIndex::search" if threads == 1 else "HnswIndex::search_batch"
        self._lock.acquire_read()
        try:
            out, count = self._core.search_arrays(queries, want, breadth, where=where)
        finally:
            self._lock.release_read()
        cols = np.arange(want)[None, :]
        found = cols < count[:, None]
        ids[found] = out["id"][:, :want][found].astype(np.int64)
        distances[found] = out["distance"][:, :want][found]
        return ids, distances

    def save(self, path) -> None:
        self._lock.acquire_read()
        try:
            self._core.save(path)
        finally:
            self._lock.release_read()

    @staticmethod
    def load(path, *, device: int = 0) -> "Index":
        return Index("l2", 1, _core=HnswIndex.load(path, device=device))

    def __repr__(self) -> str:
        return (f"<vecdex.Index space='{self.space}' dim={self.dim} M={self.M} "
                f"ef_construction={self.ef_construction} ef={self._ef} len={len(self)}>")
