"""Test setup: the `gpu` marker, the repository on sys.path, and built native libraries.

Tests load what is built and build only when nothing is (a machine that runs the GPU tests
may receive the libraries prebuilt, with no compiler)."""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")
    lib = ROOT / "hnsw_vector_search_amd" / "libvdxgpu.so"
    if not lib.exists():
        from hnsw_vector_search_amd import build
        build.build(force=False)
    oracle_lib = ROOT / "oracle" / "_build" / "libvdx_oracle.so"
    if not oracle_lib.exists():
        import subprocess
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


@pytest.fixture(scope="session")
def tmp_index_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("indexes")
