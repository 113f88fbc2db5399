"""MI355X-native query path of vecdex (HNSW search and exact brute force) behind vecdex's
own interfaces. See index.py for the classes and include/vdxgpu.h for the C ABI.

The classes load libvdxgpu.so on first use, so `python -m hnsw_vector_search_amd.build` runs
before the library exists or while it is stale."""

__version__ = "0.1.0"
__all__ = ["BruteForceIndex", "GraphReport", "HnswIndex", "Index", "Metric", "Neighbor",
           "__version__"]


def __getattr__(name):
    if name in __all__:
        from . import index
        return getattr(index, name)
    raise AttributeError(name)
