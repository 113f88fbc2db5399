"""The restated CPU graph build and FORMAT v1, pinned by vecdex's golden vectors.

The build is the product's (host C++); searches here are the oracle's, so the golden
fingerprints pin the build and the oracle's search together."""

import struct

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import datasets
from oracle import oracle as o

GREEDY_FP = 6133164406513967666   # tests/hnsw_parallel_test.cpp:164
TOP10_FP = 15666754926265556722   # tests/hnsw_parallel_test.cpp:165


def fingerprint_of(graph, queries, k, ef):
    out, cnt = o.hnsw_search(graph, queries, k, ef, threads=4)
    return o.fingerprint([(out[i]["id"][:cnt[i]], out[i]["distance"][:cnt[i]])
                          for i in range(len(queries))])


@pytest.mark.parametrize("how", ["one_by_one", "one_thread"])
def test_serial_build_reproduces_the_golden_fingerprints(how):
    base = datasets.integer_matrix(1000, 16, 1)
    queries = datasets.integer_matrix(200, 16, 2)
    index = vdx.HnswIndex(16, vdx.Metric.L2, M=8, ef_construction=64, seed=42)
    if how == "one_by_one":
        for row in base:
            index.add(row)
    else:
        index.add(base, num_threads=1)
    g = index.graph_arrays()
    assert fingerprint_of(g, queries, 1, 1) == GREEDY_FP
    assert fingerprint_of(g, queries, 10, 10) == TOP10_FP
    report = index.inspect()
    assert report.max_level == 3
    assert report.mean_degree0 * 1000 == 11157.0
    assert report.violations == [] and report.reachable == 1000


def test_a_planted_mistake_changes_the_fingerprint():
    """A graph one link off (as a mis-restated build would give) no longer matches."""
    base = datasets.integer_matrix(1000, 16, 1)
    queries = datasets.integer_matrix(200, 16, 2)
    index = vdx.HnswIndex(16, vdx.Metric.L2, M=8, ef_construction=64, seed=42)
    index.add(base)
    g = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in index.graph_arrays().items()}
    b0 = 1 + 2 * 8
    for node in range(1000):
        if g["layer0"][node * b0] > 2:
            g["layer0"][node * b0] -= 1  # drop the last link of every node with several
    assert fingerprint_of(g, queries, 10, 10) != TOP10_FP


def test_graph_invariants_and_reachability():
    rng = np.random.default_rng(1)
    base = rng.uniform(-1, 1, (2000, 16)).astype(np.float32)
    index = vdx.HnswIndex(16, vdx.Metric.L2, M=8, ef_construction=64)
    index.add(base)
    r = index.inspect()
    assert r.violations == [] and r.reachable == 2000 and r.max_level >= 1


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_exact_when_ef_covers_every_node(metric):
    # tests/hnsw_test.cpp:101-111, for every metric
    rng = np.random.default_rng(1)
    base = rng.uniform(-1, 1, (300, 8)).astype(np.float32)
    queries = rng.uniform(-1, 1, (20, 8)).astype(np.float32)
    index = vdx.HnswIndex(8, metric, M=8, ef_construction=50)
    index.add(base)
    assert index.inspect().reachable == 300
    g = index.graph_arrays()
    prepared = o.normalize(queries) if metric == 2 else queries
    out, cnt = o.hnsw_search(g, prepared, 10, 300)
    exact = o.bruteforce_f32(g["vectors"], prepared, 10, metric)
    np.testing.assert_array_equal(out, exact)


def recall(graph, queries, truth, k, ef):
    out, _ = o.hnsw_search(graph, queries, k, ef)
    return o.recall_at_k(out["id"], truth, k)


def test_parallel_build_is_healthy_and_as_good_as_serial():
    # tests/hnsw_parallel_test.cpp:117-136, 175-178
    base = datasets.integer_matrix(2000, 16, 1)
    queries = datasets.integer_matrix(1000, 16, 2)
    truth, _ = o.bruteforce_f64(base, queries, 10)
    serial = vdx.HnswIndex(16, vdx.Metric.L2, M=16, ef_construction=64, seed=42)
    serial.add(base, 1)
    r_serial = recall(serial.graph_arrays(), queries, truth, 10, 16)
    for threads in (2, 4, 8):
        par = vdx.HnswIndex(16, vdx.Metric.L2, M=16, ef_construction=64, seed=42)
        par.add(base, threads)
        rep = par.inspect()
        assert rep.violations == [] and rep.reachable >= 1998
        assert rep.max_level == serial.inspect().max_level
        assert recall(par.graph_arrays(), queries, truth, 10, 16) >= r_serial - 0.005


def test_rejects_bad_parameters():
    # tests/hnsw_test.cpp:70-77
    with pytest.raises(ValueError):
        vdx.HnswIndex(0)
    with pytest.raises(ValueError, match="M must be at least 2"):
        vdx.HnswIndex(4, vdx.Metric.L2, M=1)
    with pytest.raises(ValueError, match="M must be at most 10000"):
        vdx.HnswIndex(4, vdx.Metric.L2, M=10001)
    with pytest.raises(ValueError, match="ef_construction must be positive"):
        vdx.HnswIndex(4, vdx.Metric.L2, M=16, ef_construction=0)
    vdx.HnswIndex(4, vdx.Metric.L2, M=10000)


def test_add_rejects_bad_vectors_and_changes_nothing():
    index = vdx.HnswIndex(2)
    bad = np.zeros((3, 2), np.float32)
    bad[2, 1] = np.nan
    with pytest.raises(ValueError, match="HnswIndex::add: vectors must not contain NaN or inf"):
        index.add(bad)
    assert index.size() == 0
    with pytest.raises(ValueError, match="got dimension 3, index has 2"):
        index.add(np.zeros(3, np.float32))
    cos = vdx.HnswIndex(2, vdx.Metric.Cosine)
    with pytest.raises(ValueError, match="vector 1 has zero length"):
        cos.add(np.array([[1, 0], [0, 0]], np.float32))


# ---- FORMAT v1 ----------------------------------------------------------------------------

FORMAT_EXAMPLE_HEADER = bytes.fromhex(
    "8956 4458 484e 5357 0100 0000 0000 0000"
    "0200 0000 0000 0000 0200 0000 0000 0000"
    "0800 0000 0000 0000 2a00 0000 0000 0000"
    "0c00 0000 0000 0000 0300 0000 0000 0000"
    "0400 0000 0000 0000 0a00 0000 0000 0000".replace(" ", ""))


def test_format_example_header(tmp_path):
    """docs/FORMAT.md:52-64: 12 nodes, dim 2, M 2, ef_construction 8, seed 42, L2 -> max level 3,
    entry point 4, U = 10, 556 bytes. The levels come from the seeded generator only."""
    index = vdx.HnswIndex(2, vdx.Metric.L2, M=2, ef_construction=8, seed=42)
    index.add(np.random.default_rng(0).uniform(-1, 1, (12, 2)).astype(np.float32))
    path = tmp_path / "tiny.vdx"
    index.save(path)
    data = path.read_bytes()
    assert len(data) == 556
    assert data[:80] == FORMAT_EXAMPLE_HEADER
    assert struct.unpack("<Q", data[-8:])[0] == o.xxh64(data[:-8])


def test_file_follows_the_documented_layout(tmp_path):
    # tests/hnsw_io_test.cpp:375-417
    rng = np.random.default_rng(1)
    vectors = rng.uniform(-1, 1, (40, 3)).astype(np.float32)
    index = vdx.HnswIndex(3, vdx.Metric.L2, M=5, ef_construction=17, seed=1234)
    index.add(vectors)
    path = tmp_path / "x.vdx"
    index.save(path)
    data = path.read_bytes()
    magic, version, metric, dim, M, efc, seed, n, max_level, entry, U = struct.unpack(
        "<8sII8Q", data[:80])
    assert (magic, version, metric, dim, M, efc, seed, n) == (b"\x89VDXHNSW", 1, 0, 3, 5, 17, 1234, 40)
    levels = np.frombuffer(data, np.uint8, n, 80)
    assert levels[entry] == max_level == index.inspect().max_level and levels.sum() == U
    assert len(data) == 88 + n + 4 * n * (1 + 2 * M) + 4 * U * (1 + M) + 4 * n * dim
    vec_at = 80 + n + 4 * n * (1 + 2 * M) + 4 * U * (1 + M)
    np.testing.assert_array_equal(np.frombuffer(data, np.float32, n * dim, vec_at).reshape(n, dim),
                                  vectors)
    for code, metric in ((0, vdx.Metric.L2), (1, vdx.Metric.InnerProduct), (2, vdx.Metric.Cosine)):
        vdx.HnswIndex(2, metric).save(tmp_path / "m.vdx")
        assert struct.unpack("<I", (tmp_path / "m.vdx").read_bytes()[12:16])[0] == code


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_round_trip_is_identical(tmp_path, metric):
    # tests/hnsw_io_test.cpp:271-303, 337-351
    rng = np.random.default_rng(2)
    base = rng.uniform(-1, 1, (500, 8)).astype(np.float32)
    extra = rng.uniform(-1, 1, (50, 8)).astype(np.float32)
    queries = rng.uniform(-1, 1, (30, 8)).astype(np.float32)
    index = vdx.HnswIndex(8, metric, M=8, ef_construction=40)
    index.add(base)
    index.save(tmp_path / "a.vdx")
    loaded = vdx.HnswIndex.load(tmp_path / "a.vdx")
    loaded.save(tmp_path / "b.vdx")
    assert (tmp_path / "a.vdx").read_bytes() == (tmp_path / "b.vdx").read_bytes()
    prepared = o.normalize(queries) if metric == 2 else queries
    for ef in (1, 10, 40, 200):
        a, _ = o.hnsw_search(index.graph_arrays(), prepared, 10, ef)
        b, _ = o.hnsw_search(loaded.graph_arrays(), prepared, 10, ef)
        np.testing.assert_array_equal(a, b)
    index.add(extra)
    loaded.add(extra)
    index.save(tmp_path / "c.vdx")
    loaded.save(tmp_path / "d.vdx")
    assert (tmp_path / "c.vdx").read_bytes() == (tmp_path / "d.vdx").read_bytes()


def tiny_file(tmp_path):
    index = vdx.HnswIndex(2, vdx.Metric.L2, M=2, ef_construction=8, seed=42)
    index.add(np.random.default_rng(0).uniform(-1, 1, (12, 2)).astype(np.float32))
    index.save(tmp_path / "tiny.vdx")
    return (tmp_path / "tiny.vdx").read_bytes()


def test_every_truncation_is_rejected(tmp_path):
    # tests/hnsw_io_test.cpp:422-431
    data = tiny_file(tmp_path)
    cut = tmp_path / "cut.vdx"
    for length in range(0, len(data)):
        cut.write_bytes(data[:length])
        reason = "too small" if length < 88 else "truncated or corrupted"
        with pytest.raises(RuntimeError, match=reason) as err:
            vdx.HnswIndex.load(cut)
        assert str(err.value).startswith(str(cut))


def test_every_flipped_bit_is_rejected(tmp_path):
    # tests/hnsw_io_test.cpp:434-445
    data = bytearray(tiny_file(tmp_path))
    path = tmp_path / "flip.vdx"
    for offset in range(len(data)):
        for bit in range(8):
            data[offset] ^= 1 << bit
            path.write_bytes(bytes(data))
            with pytest.raises(RuntimeError):
                vdx.HnswIndex.load(path)
            data[offset] ^= 1 << bit


def reseal(data: bytearray) -> bytes:
    data[-8:] = struct.pack("<Q", o.xxh64(bytes(data[:-8])))
    return bytes(data)


def test_hostile_headers_and_graphs_are_rejected(tmp_path):
    # docs/FORMAT.md:104-117; tests/hnsw_io_test.cpp:487-631
    data = tiny_file(tmp_path)
    path = tmp_path / "bad.vdx"
    cases = [
        (0, b"\x89VDXHNSX", "wrong magic bytes"),
        (8, struct.pack("<I", 3), "format version 3 is not supported"),
        # version 2 is this library's aligned format: a v1 body cannot pass for one
        (8, struct.pack("<I", 2), "too small for a version 2 vecdex index"),
        (12, struct.pack("<I", 3), "unknown metric code 3"),
        (24, struct.pack("<Q", 1), "M is 1, below the minimum of 2"),
        (32, struct.pack("<Q", 0), "ef_construction is 0"),
        (64, struct.pack("<Q", 12), "entry point 12 is not one of the 12 nodes"),
        (56, struct.pack("<Q", 256), "max level 256 is above 255"),
    ]
    for offset, patch, message in cases:
        bad = bytearray(data)
        bad[offset:offset + len(patch)] = patch
        path.write_bytes(reseal(bad))
        with pytest.raises(RuntimeError, match=message):
            vdx.HnswIndex.load(path)
    # a link to a missing node, resealed so the checksum passes
    n, M = 12, 2
    bad = bytearray(data)
    first_block = 80 + n
    bad[first_block:first_block + 8] = struct.pack("<II", 1, 99)
    path.write_bytes(reseal(bad))
    with pytest.raises(RuntimeError, match="links to node 99, but there are only 12 nodes"):
        vdx.HnswIndex.load(path)
    bad = bytearray(data)
    bad[first_block:first_block + 8] = struct.pack("<II", 1, 0)
    path.write_bytes(reseal(bad))
    with pytest.raises(RuntimeError, match="node 0 links to itself on layer 0"):
        vdx.HnswIndex.load(path)
    bad = bytearray(data)
    bad[first_block:first_block + 4] = struct.pack("<I", 2 * M + 1)
    path.write_bytes(reseal(bad))
    with pytest.raises(RuntimeError, match="node 0 has 5 links on layer 0 \\(max 4\\)"):
        vdx.HnswIndex.load(path)


def test_missing_file_names_the_path(tmp_path):
    with pytest.raises(RuntimeError) as err:
        vdx.HnswIndex.load(tmp_path / "nope.vdx")
    assert str(err.value).startswith(str(tmp_path / "nope.vdx"))
