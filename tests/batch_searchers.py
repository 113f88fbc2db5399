"""Test helpers for the batched add: a vdx_batch_searcher over the oracle's search."""

import ctypes

import numpy as np

from hnsw_vector_search_amd import _native as nat
from oracle import oracle as o


def arrays(v: nat.GraphView) -> dict:
    def arr(p, count, dtype):
        if count == 0 or not p:
            return np.zeros(0, dtype=dtype)
        buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(p)
        return np.frombuffer(buf, dtype=dtype, count=count)
    return {"metric": v.metric, "dim": v.dim, "n": v.n, "M": v.M, "max_level": v.max_level,
            "entry_point": v.entry_point, "upper_blocks": v.upper_blocks,
            "levels": arr(v.levels, v.n, np.uint8),
            "layer0": arr(v.layer0, v.n * (1 + 2 * v.M), np.uint32),
            "upper": arr(v.upper, v.upper_blocks * (1 + v.M), np.uint32),
            "vectors": arr(v.vectors, v.n * v.dim, np.float32).reshape(v.n, v.dim)}


def apply_changes(copy: dict | None, live: dict, changed: np.ndarray) -> dict:
    """What vdx_gpu_hnsw_update does to the device copy, on numpy copies: the new nodes are
    appended from the live graph and only the listed old nodes' blocks are rewritten; every other
    block keeps the copy's contents. A node missing from the list stays stale, as on the GPU."""
    fresh = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in live.items()}
    if copy is None:
        return fresh
    old_n, M = copy["n"], live["M"]
    b0, bu = 1 + 2 * M, 1 + M
    layer0 = fresh["layer0"]
    layer0[: old_n * b0] = copy["layer0"]
    upper = fresh["upper"]
    upper[: copy["upper"].size] = copy["upper"]
    offsets = np.concatenate([[0], np.cumsum(live["levels"].astype(np.int64))])
    for i in changed[changed < old_n].astype(np.int64):
        layer0[i * b0:(i + 1) * b0] = live["layer0"][i * b0:(i + 1) * b0]
        lo, hi = offsets[i] * bu, offsets[i + 1] * bu
        upper[lo:hi] = live["upper"][lo:hi]
    return fresh


class OracleSearcher:
    """A vdx_batch_searcher whose copy is the synced view and whose searches are the oracle's."""

    def __init__(self, fail_search: int = 0, graph: dict | None = None, upper_layers: bool = False):
        self.graph = graph
        self.syncs = 0
        self.batches = []
        self.entries = []
        self.layers = []
        self.changed = []

        def sync(ctx, view, changed, count):
            ids = np.ctypeslib.as_array(changed, (count,)).copy() if count else np.zeros(0, np.uint32)
            self.graph = apply_changes(self.graph, arrays(view.contents), ids)
            self.changed.append(ids)
            self.syncs += 1
            return 0

        def search(ctx, nodes, entries, count, layer, ef, m, out, counts):
            if m != 0 or (layer != 0 and not upper_layers):
                return 2  # this searcher returns results; the graph makes the choice
            if fail_search:
                return fail_search
            ids = np.ctypeslib.as_array(nodes, (count,)).copy()
            ent = np.ctypeslib.as_array(entries, (count,)).copy()
            self.batches.append(ids)
            self.entries.append(ent)
            g = self.graph
            self.layers.append(layer)
            found, cnt = o.hnsw_search(g, g["vectors"][ids], ef, ef, entries=ent, layer=layer)
            ctypes.memmove(out, found.ctypes.data, found.nbytes)
            ctypes.memmove(counts, cnt.ctypes.data, cnt.nbytes)
            return 0

        self._keep = (nat.SYNC_FN(sync), nat.SEARCH_FN(search))
        self.c = nat.BatchSearcher(None, *self._keep, nat.SEARCHER_UPPER_LAYERS if upper_layers else 0)


def add_batched(index, data, threads, plan, searcher):
    data = np.ascontiguousarray(data, dtype=np.float32)
    version = index._graph.params()[2]
    return nat.lib.vdx_host_graph_add_batched(index._graph.handle, nat.ptr(data), data.shape[0],
                                              data.shape[1], threads, ctypes.byref(searcher.c),
                                              ctypes.byref(plan), version, None)


def sections(index):
    g = index.graph_arrays()
    return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in g.items()}
