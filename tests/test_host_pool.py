"""The host thread pool that stages queries and unpacks results in the search entry points
(csrc/host/host_pool.cpp), driven from a small C++ program built here with g++: every item of
every batch runs exactly once; a failing batch rethrows the lowest failing item's exception (what
a sequential loop throws first) and runs no item after a failure it has seen; a second caller
while the pool is busy runs its items itself; batches come back to back (workers still spinning)
and after pauses longer than the spin (workers asleep). The same program again under
ThreadSanitizer, where the runtime is installed."""

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HOST = ROOT / "hnsw_vector_search_amd" / "csrc" / "host"

DRIVER = r"""
#include "host_pool.hpp"
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

static int fail(const char* what, long a, long b) {
    std::fprintf(stderr, "FAIL %s %ld %ld\n", what, a, b);
    return 1;
}

static int batches(vdx::HostPool& pool, unsigned seed, int rounds) {
    std::mt19937 rng(seed);
    std::vector<std::atomic<int>> hits(4096);
    for (int r = 0; r < rounds; ++r) {
        const std::size_t count = 1 + rng() % 600;
        for (std::size_t i = 0; i < count; ++i) hits[i].store(0);
        const int mode = static_cast<int>(rng() % 4);
        std::size_t bad = count;  // lowest failing item
        std::vector<char> throws(count, 0);
        if (mode == 0) {
            for (int j = 0; j < 3; ++j) {
                const std::size_t i = rng() % count;
                throws[i] = 1;
                bad = std::min(bad, i);
            }
        }
        long caught = -1;
        try {
            pool.run(count, [&](std::size_t i) {
                hits[i].fetch_add(1);
                if (throws[i]) throw std::runtime_error(std::to_string(i));
                if ((i & 63) == 0) std::this_thread::yield();
            });
        } catch (const std::runtime_error& e) {
            caught = std::atol(e.what());
        }
        if (bad < count) {
            if (caught != static_cast<long>(bad)) return fail("lowest failure", caught, static_cast<long>(bad));
            for (std::size_t i = 0; i <= bad; ++i)
                if (hits[i].load() != 1) return fail("item below a failure", static_cast<long>(i), hits[i].load());
        } else {
            if (caught != -1) return fail("unexpected failure", caught, 0);
            for (std::size_t i = 0; i < count; ++i)
                if (hits[i].load() != 1) return fail("item count", static_cast<long>(i), hits[i].load());
        }
        if (mode == 1 && r % 16 == 0) {
            std::this_thread::sleep_for(vdx::HostPool::kSpin * 3);  // the workers fall asleep
        }
    }
    return 0;
}

int main() {
    vdx::HostPool pool(8);
    if (pool.threads() != 8) return fail("threads", static_cast<long>(pool.threads()), 8);
    if (batches(pool, 1, 1500)) return 1;
    // two callers at once: one runs on the pool, the other inline; both see every item once
    int rc[2] = {0, 0};
    std::thread a([&] { rc[0] = batches(pool, 2, 600); });
    std::thread b([&] { rc[1] = batches(pool, 3, 600); });
    a.join();
    b.join();
    if (rc[0] || rc[1]) return 1;
    vdx::HostPool single(1);  // no workers: the caller runs everything
    if (batches(single, 4, 200)) return 1;
    std::puts("ok");
    return 0;
}
"""


def _build_and_run(tmp_path, flags, name):
    src = tmp_path / "pool_driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / name
    cmd = ["g++", "-std=c++20", "-g", *flags, f"-I{HOST}", str(src), str(HOST / "host_pool.cpp"),
           "-o", str(exe), "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert proc.returncode == 0 and proc.stdout.strip() == "ok", proc.stdout + proc.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_every_item_once_and_the_lowest_failure(tmp_path):
    _build_and_run(tmp_path, ["-O2", "-Wall", "-Wextra"], "pool")


def _has_tsan() -> bool:
    if shutil.which("gcc") is None:
        return False
    path = subprocess.run(["gcc", "-print-file-name=libtsan.so"], capture_output=True, text=True).stdout.strip()
    return Path(path).is_absolute() and Path(path).exists()


@pytest.mark.skipif(not _has_tsan(), reason="needs g++ and the ThreadSanitizer runtime")
def test_no_data_race_under_thread_sanitizer(tmp_path):
    _build_and_run(tmp_path, ["-O1", "-fsanitize=thread"], "pool_tsan")
