"""vecdex.Index semantics that need no device: input rules and messages, ids, properties,
save/load (python/tests/test_index.py:166-244, 73-100, 146-160)."""

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx

DIM = 16


def vectors(n, seed=0):
    return np.random.default_rng(seed).standard_normal((n, DIM), dtype=np.float32)


def test_add_items_returns_consecutive_int64_ids():
    index = vdx.Index("l2", DIM)
    first = index.add_items(vectors(10))
    second = index.add_items(vectors(5, 1))
    assert first.dtype == np.int64 and first.shape == (10,)
    np.testing.assert_array_equal(first, np.arange(10))
    np.testing.assert_array_equal(second, np.arange(10, 15))
    assert len(index) == 15


def test_properties_and_repr():
    index = vdx.Index("ip", 8, M=12, ef_construction=50)
    index.add_items(np.random.default_rng(0).standard_normal((3, 8), dtype=np.float32))
    assert (index.space, index.dim, index.M, index.ef_construction) == ("ip", 8, 12, 50)
    assert index.memory_bytes() >= 3 * 8 * 4
    assert repr(index) == "<vecdex.Index space='ip' dim=8 M=12 ef_construction=50 ef=64 len=3>"


def test_ef_property():
    index = vdx.Index("l2", DIM)
    assert index.ef == 64
    index.set_ef(200)
    assert index.ef == 200
    index.ef = 10
    assert index.ef == 10
    with pytest.raises(ValueError, match="ef must be positive, got 0"):
        index.ef = 0


@pytest.mark.parametrize("dtype", [np.float64, np.float16, np.int32, ">f4"])
def test_other_dtypes_are_rejected_with_type_error(dtype):
    index = vdx.Index("l2", DIM)
    with pytest.raises(TypeError, match=r"must be float32.*astype"):
        index.add_items(vectors(4).astype(dtype))
    with pytest.raises(TypeError, match=r"must be float32.*astype"):
        index.knn_query(vectors(4).astype(dtype))


def test_shape_contiguity_and_alignment_rules():
    index = vdx.Index("l2", DIM)
    with pytest.raises(ValueError, match="must be 1-D .* got 3-D"):
        index.add_items(np.zeros((2, 2, DIM), np.float32))
    with pytest.raises(ValueError, match="vectors have dimension 8, the index has 16"):
        index.add_items(np.zeros((2, 8), np.float32))
    with pytest.raises(ValueError, match="C-contiguous"):
        index.add_items(np.zeros((DIM, 4), np.float32).T)
    raw = np.zeros(4 * DIM + 1, dtype=np.uint8)[1:].view(np.float32)
    with pytest.raises(ValueError, match="not aligned"):
        index.add_items(raw)


def test_bad_arguments():
    with pytest.raises(ValueError, match="space must be 'l2', 'ip' or 'cosine', got 'L1'"):
        vdx.Index("L1", DIM)
    with pytest.raises(ValueError, match="dim must be positive"):
        vdx.Index("l2", 0)
    index = vdx.Index("l2", DIM)
    with pytest.raises(ValueError, match="num_threads must be positive, or 0"):
        index.add_items(vectors(2), num_threads=-1)
    with pytest.raises(ValueError, match="k must be positive, got 0"):
        index.knn_query(vectors(2), k=0)


def test_rejected_add_adds_nothing():
    index = vdx.Index("cosine", DIM)
    data = vectors(5)
    data[3] = 0
    with pytest.raises(ValueError, match="vector 3 has zero length"):
        index.add_items(data)
    assert len(index) == 0


def test_save_load_keeps_parameters(tmp_path):
    index = vdx.Index("cosine", DIM, M=8, ef_construction=40, seed=7)
    index.add_items(vectors(100))
    index.ef = 99
    index.save(tmp_path / "i.vdx")
    loaded = vdx.Index.load(str(tmp_path / "i.vdx"))
    assert (loaded.space, loaded.dim, loaded.M, loaded.ef_construction, len(loaded)) == \
        ("cosine", DIM, 8, 40, 100)
    assert loaded.ef == 64  # ef is not saved
    with pytest.raises(RuntimeError, match="nope"):
        vdx.Index.load(tmp_path / "nope.vdx")


def test_reader_writer_lock_lets_a_writer_in_between_back_to_back_readers():
    """Three threads re-taking the read lock back to back (searches in a loop) must not starve
    a writer (add_items): once it waits, new readers wait behind it."""
    import threading
    import time
    from hnsw_vector_search_amd.index import _RWLock

    lock = _RWLock()
    stop = threading.Event()
    reads = [0, 0, 0]

    def reader(i):
        while not stop.is_set():
            lock.acquire_read()
            try:
                time.sleep(0.0005)
                reads[i] += 1
            finally:
                lock.release_read()

    threads = [threading.Thread(target=reader, args=(i,)) for i in range(3)]
    for t in threads:
        t.start()
    try:
        time.sleep(0.05)
        writes = 0
        start = time.perf_counter()
        while writes < 20 and time.perf_counter() - start < 10:
            lock.acquire_write()
            writes += 1
            lock.release_write()
    finally:
        stop.set()
        for t in threads:
            t.join()
    assert writes == 20
    assert min(reads) > 0
