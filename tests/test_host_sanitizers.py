"""The host code's CPU tests again under AddressSanitizer and UndefinedBehaviorSanitizer.

vecdex runs its loader under ASan + UBSan (CMakeLists.txt:20-30, tests/load_fuzzer.cpp). Here
the library's host sources (FORMAT v1/v2 parsers, TEXMEX reader, graph build, validation,
recall) are rebuilt with both sanitizers (`build.py --sanitize`, linked with the regular objects
of the HIP sources) and the CPU test files that drive them — every truncation and flipped bit of
small index files, hostile headers, malformed .fvecs/.ivecs — run against that build in a child
process with the sanitizer runtimes preloaded; any report aborts the child."""

import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
FILES = ["test_format_v2.py", "test_graph_build.py", "test_vecs_eval.py", "test_recall.py",
         "test_graph_changes.py", "test_batched_build.py", "test_index_cpu.py", "test_capi.py"]


def _runtime(name: str) -> str | None:
    if shutil.which("gcc") is None:
        return None
    path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.skipif(shutil.which("g++") is None or _runtime("libasan.so") is None,
                    reason="needs g++ and its sanitizer runtimes (the build machine)")
def test_host_code_is_clean_under_asan_and_ubsan():
    sys.path.insert(0, str(ROOT))
    from hnsw_vector_search_amd import build
    build.build_sanitized()
    env = dict(os.environ)
    env.update(VDX_LIBRARY="sanitized",
               LD_PRELOAD=f"{_runtime('libasan.so')} {_runtime('libubsan.so')}",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    proc = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                           *(str(ROOT / "tests" / f) for f in FILES)],
                          cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    report = (proc.stdout + proc.stderr)[-4000:]
    assert proc.returncode == 0, report
    assert "runtime error" not in proc.stderr and "AddressSanitizer" not in proc.stderr, report
