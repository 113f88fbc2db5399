"""Builds libvdxgpu.so in-tree: host C++ with g++, HIP kernels for gfx950 with hipcc.

    python -m hnsw_vector_search_amd.build            # incremental
    python -m hnsw_vector_search_amd.build --force    # everything
    python -m hnsw_vector_search_amd.build --sanitize # host code under ASan + UBSan (below)

Host sources are compiled without -mavx2 (the AVX2 kernels carry a function target attribute
and run only after a CPU check) and with -ffp-contract=off, as are the HIP sources: every
multiply and add rounds as written, which the exact summation order depends on. HIP code is
also built with packed-FP32 instructions disabled.
"""

from __future__ import annotations

import argparse
import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libvdxgpu.so"
SAN_BUILD = PKG / "_build_san"
SAN_LIB = SAN_BUILD / "libvdxgpu.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("VDX_OFFLOAD_ARCH", "gfx950")

HOST_FLAGS = ["-std=c++20", "-O3", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
              "-fvisibility=hidden"]
HIP_FLAGS = ["-std=c++20", "-O3", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fvisibility=hidden", "-Wno-unused-command-line-argument"]


def _sources():
    host = sorted((CSRC / "host").glob("*.cpp"))
    dev = sorted((CSRC / "device").glob("*.hip"))
    return host, dev


def _headers():
    return sorted(CSRC.rglob("*.hpp")) + [PKG.parent / "include" / "vdxgpu.h"]


def _obj(src: Path) -> Path:
    return BUILD / (src.parent.name + "_" + src.stem + ".o")


def _compile(src: Path, force: bool, newest_header: float) -> Path:
    obj = _obj(src)
    if (not force and obj.exists()
            and obj.stat().st_mtime >= max(src.stat().st_mtime, newest_header)):
        return obj
    if src.suffix == ".hip":
        cmd = [str(ROCM / "bin" / "hipcc"), *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *HOST_FLAGS, "-c", str(src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{proc.stdout}{proc.stderr}")
    return obj


def build(force: bool = True, jobs: int = 8, verbose: bool = False) -> Path:
    """Compiles every source (all of them when force) and links libvdxgpu.so."""
    BUILD.mkdir(exist_ok=True)
    host, dev = _sources()
    newest_header = max(h.stat().st_mtime for h in _headers())
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(lambda s: _compile(s, force, newest_header), host + dev))
    if (force or not LIB.exists()
            or LIB.stat().st_mtime < max(o.stat().st_mtime for o in objs)):
        tmp = LIB.with_suffix(".so.tmp")
        cmd = [str(ROCM / "bin" / "hipcc"), f"--offload-arch={ARCH}", "-shared", "-fPIC",
               "-o", str(tmp), *map(str, objs), "-lpthread"]
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            raise RuntimeError(f"link failed: {' '.join(cmd)}\n{proc.stdout}{proc.stderr}")
        os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


SAN_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
             "-g", "-O1"]


def build_sanitized(jobs: int = 8, verbose: bool = False) -> Path:
    """The same library with every host source (the file parsers, graph build, validation) built
    by g++ under AddressSanitizer and UndefinedBehaviorSanitizer, linked with the regular
    objects of the HIP sources: a CPU-side check of the host code (load it with the sanitizer
    runtime preloaded). Written to _build_san/, never over the product library."""
    build(force=False, jobs=jobs)
    SAN_BUILD.mkdir(exist_ok=True)
    host, dev = _sources()

    def compile_host(src: Path) -> Path:
        obj = SAN_BUILD / (src.stem + ".o")
        cmd = ["g++", *HOST_FLAGS, *SAN_FLAGS, "-c", str(src), "-o", str(obj)]
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{proc.stdout}{proc.stderr}")
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(compile_host, host))
    cmd = ["g++", "-shared", "-fPIC", *SAN_FLAGS, "-o", str(SAN_LIB), *map(str, objs),
           *(str(_obj(d)) for d in dev), f"-L{ROCM / 'lib'}", "-lamdhip64", "-lpthread",
           f"-Wl,-rpath,{ROCM / 'lib'}"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{proc.stdout}{proc.stderr}")
    if verbose:
        print(f"built {SAN_LIB}")
    return SAN_LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("--sanitize", action="store_true",
                    help="also build _build_san/libvdxgpu.so with host code under ASan + UBSan")
    ap.add_argument("-j", "--jobs", type=int, default=8)
    args = ap.parse_args(argv)
    build(force=args.force, jobs=args.jobs, verbose=True)
    if args.sanitize:
        build_sanitized(jobs=args.jobs, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
