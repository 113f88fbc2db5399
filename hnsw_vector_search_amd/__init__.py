"""MI355X-native query path of vecdex (HNSW search and exact brute force) behind vecdex's
own interfaces. See index.py for the classes, recall.py for the recall measures and
include/vdxgpu.h for the C ABI.

The classes load libvdxgpu.so on first use, so `python -m hnsw_vector_search_amd.build` runs
before the library exists or while it is stale."""

__version__ = "0.1.0"
_INDEX = ("BruteForceIndex", "GraphReport", "HnswDeviceIndex", "HnswIndex", "Index", "Metric", "Neighbor",
          "SearchBatch")
_RECALL = ("recall_at_k", "threshold_recall_at_k", "mean_recall_at_k", "mean_threshold_recall_at_k")
__all__ = [*_INDEX, *_RECALL, "__version__"]


def __getattr__(name):
    if name in _INDEX:
        from . import index
        return getattr(index, name)
    if name in _RECALL:
        from . import recall
        return getattr(recall, name)
    raise AttributeError(name)
