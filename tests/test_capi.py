"""The C ABI library loads and exports every entry point include/vdxgpu.h declares (no device
calls here)."""

import re
import subprocess
from pathlib import Path

from hnsw_vector_search_amd import _native as nat

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include" / "vdxgpu.h").read_text()
    return sorted(set(re.findall(r"^VDX_API [^;(]*?\b(vdx_\w+)\(", text, flags=re.M)))


def test_header_declares_the_boundary():
    names = declared()
    for required in ("vdx_gpu_hnsw_search", "vdx_gpu_hnsw_search_device", "vdx_gpu_bf_search",
                     "vdx_gpu_bf_search_device", "vdx_gpu_hnsw_upload", "vdx_host_graph_load"):
        assert required in names


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", str(nat.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (vdx_\w+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    for name in declared():
        assert hasattr(nat.lib, name)


def test_binding_covers_every_declared_symbol():
    assert set(declared()) == set(nat.EXPORTED)


def test_errors_map_to_python_exception_types():
    import ctypes
    h = ctypes.c_void_p()
    assert nat.lib.vdx_host_graph_create(0, 0, 16, 200, 42, ctypes.byref(h)) == nat.VDX_INVALID_ARGUMENT
    assert b"dim must be positive" in nat.lib.vdx_last_error()
    assert nat.lib.vdx_host_graph_load(b"/nonexistent/x.vdx", ctypes.byref(h)) == nat.VDX_RUNTIME_ERROR
    assert nat.lib.vdx_last_error().startswith(b"/nonexistent/x.vdx")


def test_host_distance_is_the_exact_order():
    import numpy as np
    from oracle import oracle as o
    rng = np.random.default_rng(0)
    for dim in (1, 5, 31, 100, 128, 287):
        a = rng.standard_normal(dim).astype(np.float32)
        b = rng.standard_normal(dim).astype(np.float32)
        assert nat.lib.vdx_host_distance(0, a.ctypes.data, b.ctypes.data, dim) == \
            np.float32(o.distance("l2_avx2order", a, b))
        assert nat.lib.vdx_host_distance(1, a.ctypes.data, b.ctypes.data, dim) == \
            np.float32(1) - np.float32(o.distance("ip_avx2order", a, b))
