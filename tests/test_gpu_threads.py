"""Concurrent calls on one index: the guarantees of vecdex's search (`search`/`search_batch`
are const and may run from many threads at once, include/vecdex/hnsw.hpp:58-59) and of its
Python binding (knn_query releases the GIL and runs under the index's reader-writer lock,
python/bindings.cpp:190-313), tested as vecdex tests them (tests/hnsw_test.cpp:232-259,
python/tests/test_threads.py:58-264) against the GPU path. Host-entry calls share one device
copy and borrow device workspaces from the context's pool; device-entry calls on several
streams take the index's queue-counter slots in turn."""

import ctypes
import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import _native as nat
from hnsw_vector_search_amd import datasets

pytestmark = pytest.mark.gpu

DIM = 16
SPACES = ["l2", "ip", "cosine"]


def random_vectors(n, dim=DIM, seed=0):
    return np.random.default_rng(seed).standard_normal((n, dim), dtype=np.float32)


def build(base, num_threads=1):
    index = vdx.Index("l2", DIM, M=8, ef_construction=40)
    index.add_items(base, num_threads=num_threads)
    return index


def scale_to(seconds, trial):
    start = time.perf_counter()
    trial()
    return max(1, math.ceil(seconds / max(time.perf_counter() - start, 1e-6)))


def python_ticks_during(call):
    """Ticks of another Python thread (about one per millisecond) in the middle half of call():
    a call that held the GIL would allow none there."""
    ticks, stop = [], threading.Event()

    def tick():
        last = 0.0
        while not stop.is_set():
            now = time.perf_counter()
            if now - last >= 1e-3:
                ticks.append(now)
                last = now

    worker = threading.Thread(target=tick)
    worker.start()
    try:
        start = time.perf_counter()
        call()
        end = time.perf_counter()
    finally:
        stop.set()
        worker.join()
    quarter = (end - start) / 4
    return sum(start + quarter < t < end - quarter for t in ticks)


@pytest.mark.parametrize("num_threads", [1, 4])
def test_knn_query_releases_the_gil(num_threads):
    index = build(random_vectors(2000))
    queries = random_vectors(100, seed=1)

    def query(data):
        return index.knn_query(data, k=10, ef=100, num_threads=num_threads)

    batch = np.tile(queries, (scale_to(0.2, lambda: query(queries)), 1))
    assert python_ticks_during(lambda: query(batch)) > 0


def test_concurrent_knn_queries_match_sequential_ones():
    index = build(random_vectors(3000))
    queries = random_vectors(800, seed=1)
    expected_ids, expected_distances = index.knn_query(queries, k=10, ef=50)
    with ThreadPoolExecutor(max_workers=8) as pool:
        batches = list(pool.map(lambda q: index.knn_query(q, k=10, ef=50), np.array_split(queries, 16)))
        singles = list(pool.map(lambda q: index.knn_query(q, k=10, ef=50), queries))
    for results in (batches, singles):
        np.testing.assert_array_equal(np.concatenate([ids for ids, _ in results]), expected_ids)
        np.testing.assert_array_equal(np.concatenate([d for _, d in results]), expected_distances)


def test_hnsw_concurrent_searches_match_sequential_ones():
    """tests/hnsw_test.cpp:232-259 with more threads and query sets of several sizes."""
    base, queries = datasets.gmm_base_and_queries(datasets.gmm64(), 10_000, 256)
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=16, ef_construction=100, seed=0)
    index.add(base, num_threads=8)
    sets = [(queries[i::7], 10, ef) for i, ef in zip(range(7), (16, 40, 64, 100, 200, 300, 600))]
    expected = [index.search_arrays(q, k, ef) for q, k, ef in sets]
    single = [index.search(queries[q], 10, 40) for q in range(32)]

    def work(t):
        got = [index.search_arrays(q, k, ef) for q, k, ef in sets[t % 7:] + sets[:t % 7]]
        rows = [index.search(queries[q], 10, 40) for q in range(32)]
        return t, got, rows

    with ThreadPoolExecutor(max_workers=8) as pool:
        for t, got, rows in pool.map(work, range(16)):
            order = list(range(7))[t % 7:] + list(range(7))[:t % 7]
            for j, (out, cnt) in zip(order, got):
                assert out.tobytes() == expected[j][0].tobytes() and (cnt == expected[j][1]).all(), (t, j)
            assert rows == single, t


def test_bruteforce_concurrent_searches_match_sequential_ones():
    gen = datasets.AnisotropicGaussian(100, 0.8, seed=3)
    index = vdx.BruteForceIndex(100, vdx.Metric.Cosine)
    index.add(gen.sample(30_000, 1))
    queries = gen.sample(96, 2)
    modes = ("exact", "fp32", "bf16")
    expected = {m: index.search_arrays(queries, 10, m)[0] for m in modes}
    with ThreadPoolExecutor(max_workers=6) as pool:
        for m, got in pool.map(lambda i: (modes[i % 3], index.search_arrays(queries, 10, modes[i % 3])[0]),
                               range(12)):
            assert got.tobytes() == expected[m].tobytes(), m


def test_device_entry_calls_in_flight_on_several_streams():
    """More launches in flight than the index has queue-counter slots (64), spread over four
    streams with no synchronization between them: each launch waits for its slot's previous
    user, so every output equals the same search run alone."""
    base, queries = datasets.gmm_base_and_queries(datasets.gmm64(), 20_000, 4096)
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=16, ef_construction=100, seed=0)
    index.add(base, num_threads=8)
    handle = index.gpu_handle()
    ctx = nat.Context.get(0)
    contexts = [nat.Context(0) for _ in range(4)]  # one stream each
    rng = np.random.default_rng(4)
    calls = 160
    picks = [rng.choice(len(queries), 256, replace=False) for _ in range(calls)]
    efs = [int(e) for e in rng.choice([16, 48, 96, 200], calls)]
    dqs = [nat.DeviceArray.from_numpy(queries[p], ctx) for p in picks]
    outs = [nat.DeviceArray((256, 10), nat.NEIGHBOR_DTYPE, ctx) for _ in range(calls)]
    cnts = [nat.DeviceArray(256, np.uint32, ctx) for _ in range(calls)]
    expected = [index.search_arrays(queries[p], 10, ef) for p, ef in zip(picks, efs)]
    for i in range(calls):
        nat.check(nat.lib.vdx_gpu_hnsw_search_device(handle, dqs[i].ptr, 256, None, 10, None, efs[i], 10,
                                                     outs[i].ptr, 10, cnts[i].ptr, None,
                                                     contexts[i % 4].stream))
    for c in contexts:
        nat.check(nat.lib.vdx_gpu_synchronize(c.handle))
    for i in range(calls):
        assert outs[i].numpy().tobytes() == expected[i][0].tobytes(), i
        assert (cnts[i].numpy() == expected[i][1]).all(), i
    for c in contexts:
        nat.lib.vdx_gpu_close(c.handle)
        c.handle = None


def test_device_entry_ef_bound_and_clamp_counter():
    """With per-query efs the device entry point's `ef` bounds them: ef = 0 is refused; a bound
    that is too small runs the larger queries at the launch's capacity and counts them."""
    base, queries = datasets.gmm_base_and_queries(datasets.gmm64(), 8000, 128)
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=12, ef_construction=80, seed=1)
    index.add(base, num_threads=8)
    efs = np.where(np.arange(128) % 2 == 0, 40, 300).astype(np.uint32)
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(queries, ctx)
    defs = nat.DeviceArray.from_numpy(efs, ctx)
    out = nat.DeviceArray((128, 10), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(128, np.uint32, ctx)
    handle = index.gpu_handle()
    with pytest.raises(ValueError, match="ef must bound them"):
        nat.check(nat.lib.vdx_gpu_hnsw_search_device(handle, dq.ptr, 128, None, 10, defs.ptr, 0, 10,
                                                     out.ptr, 10, cnt.ptr, None, None))
    host, host_cnt = index.search_arrays(queries, 10, efs)  # the host entry point sizes from the array
    stats = (ctypes.c_uint64 * 8)()
    opts = nat.SearchOpts()
    opts.stats = ctypes.cast(stats, ctypes.c_void_p)
    nat.check(nat.lib.vdx_gpu_hnsw_search_device(handle, dq.ptr, 128, None, 10, defs.ptr, 300, 10,
                                                 out.ptr, 10, cnt.ptr, ctypes.byref(opts), None))
    assert out.numpy().tobytes() == host.tobytes() and (cnt.numpy() == host_cnt).all()
    assert stats[6] == 0
    stats = (ctypes.c_uint64 * 8)()
    opts.stats = ctypes.cast(stats, ctypes.c_void_p)
    nat.check(nat.lib.vdx_gpu_hnsw_search_device(handle, dq.ptr, 128, None, 10, defs.ptr, 40, 10,
                                                 out.ptr, 10, cnt.ptr, ctypes.byref(opts), None))
    assert stats[6] == 64  # the ef-300 queries ran at capacity 64
    got = out.numpy()
    at64, _ = index.search_arrays(queries[1::2], 10, 64)
    assert got[1::2].tobytes() == at64.tobytes()
    assert got[0::2].tobytes() == host[0::2].tobytes()


def searches_during(index, queries, add_all):
    adding = threading.Event()
    adding.set()

    def add_then_stop():
        try:
            add_all()
        finally:
            adding.clear()

    def search_until_done():
        searches = 0
        while adding.is_set():
            ids, _ = index.knn_query(queries, k=5)
            assert ids.min() >= 0 and ids.max() < len(index)
            searches += 1
        return searches

    with ThreadPoolExecutor(max_workers=4) as pool:
        searchers = [pool.submit(search_until_done) for _ in range(3)]
        pool.submit(add_then_stop).result()
        return sum(s.result() for s in searchers)


def test_queries_during_adds_are_safe_and_leave_the_build_unchanged():
    """Each add_items updates the device copy under the write lock; searches in between see a
    consistent index, and the final graph is the single-threaded build's."""
    initial = random_vectors(500)
    batches = [random_vectors(100, seed=seed) for seed in range(1, 21)]
    queries = random_vectors(50, seed=99)
    index = build(initial)
    index.knn_query(queries, k=5)  # the device copy exists before the adds

    def add_all():
        for batch in batches:
            index.add_items(batch)

    assert searches_during(index, queries, add_all) > 0, "no search overlapped the adds"
    reference = build(np.concatenate([initial, *batches]))
    assert len(index) == len(reference) == 2500
    for got, want in zip(index.knn_query(queries, k=10), reference.knn_query(queries, k=10)):
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("space", SPACES)
def test_num_threads_never_changes_results(space):
    index = vdx.Index(space, DIM, M=8, ef_construction=40)
    index.add_items(random_vectors(3000))
    tiny = vdx.Index(space, DIM)
    tiny.add_items(random_vectors(5))
    queries = random_vectors(500, seed=1)
    for target, data, k in ((index, queries, 10), (index, queries[0], 10), (index, queries[:0], 10),
                            (tiny, queries, 8)):
        expected = target.knn_query(data, k=k, ef=50)
        for num_threads in (2, 4, 0):
            got = target.knn_query(data, k=k, ef=50, num_threads=num_threads)
            for array, want in zip(got, expected):
                np.testing.assert_array_equal(array, want)
                assert array.shape == want.shape and array.dtype == want.dtype
