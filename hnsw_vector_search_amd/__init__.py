"""MI355X-native query path of vecdex (HNSW search and exact brute force) behind vecdex's
own interfaces. See index.py for the classes and include/vdxgpu.h for the C ABI."""

from .index import BruteForceIndex, GraphReport, HnswIndex, Index, Metric, Neighbor

__version__ = "0.1.0"
__all__ = ["BruteForceIndex", "GraphReport", "HnswIndex", "Index", "Metric", "Neighbor",
           "__version__"]
