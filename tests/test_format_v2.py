"""FORMAT v2, this library's aligned and mappable index file (a new version under
docs/FORMAT.md:135-137's rule: same magic, version at offset 8). The layout is restated here
independently from the specification in format_v2.hpp (struct offsets, 4 KiB-aligned sections,
per-section XXH64 computed with the oracle's XXH64, itself pinned by xxHash's sanity vectors),
and every check the loader makes is exercised on a resealed file so that only that check fails."""

import struct

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import datasets
from oracle import oracle as o

ALIGN = 4096
NAMES = ("levels", "upper offsets", "layer-0 links", "upper links", "vectors")


def parse(data: bytes) -> dict:
    """The v2 file per its specification."""
    magic, version, metric = struct.unpack_from("<8sII", data, 0)
    dim, M, efc, seed, n, max_level, entry, U, ld, align = struct.unpack_from("<10Q", data, 16)
    table = [struct.unpack_from("<3Q", data, 96 + 24 * s) for s in range(5)]
    (header_sum,) = struct.unpack_from("<Q", data, 216)
    return dict(magic=magic, version=version, metric=metric, dim=dim, M=M, efc=efc, seed=seed, n=n,
                max_level=max_level, entry=entry, U=U, ld=ld, align=align, table=table, header_sum=header_sum)


def reseal(data: bytearray) -> bytes:
    """Recomputes every section checksum and the header checksum."""
    h = parse(bytes(data))
    for s, (off, size, _) in enumerate(h["table"]):
        struct.pack_into("<Q", data, 96 + 24 * s + 16, o.xxh64(bytes(data[off:off + size])))
    struct.pack_into("<Q", data, 216, o.xxh64(bytes(data[:216])))
    return bytes(data)


def build(dim=37, n=2000, metric=vdx.Metric.Cosine, M=6):
    index = vdx.HnswIndex(dim, metric, M=M, ef_construction=40, seed=7)
    index.add(datasets.integer_matrix(n, dim, 3) + 1, num_threads=4)
    return index


def sections(index):
    return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in index.graph_arrays().items()}


def assert_same(a, b):
    a, b = sections(a), sections(b)
    for key in ("levels", "layer0", "upper", "vectors", "max_level", "entry_point", "upper_blocks", "n"):
        assert np.array_equal(a[key], b[key]), key


def test_layout_follows_the_specification(tmp_path):
    index = build()
    path = tmp_path / "g.vdx2"
    index.save(path, version=2)
    data = path.read_bytes()
    h = parse(data)
    g = sections(index)
    n, dim, M, U = g["n"], g["dim"], g["M"], g["upper_blocks"]
    assert h["magic"] == b"\x89VDXHNSW" and h["version"] == 2 and h["align"] == ALIGN
    assert (h["metric"], h["dim"], h["M"], h["n"], h["U"]) == (2, dim, M, n, U)
    assert (h["max_level"], h["entry"]) == (g["max_level"], g["entry_point"])
    assert h["ld"] == 40 and h["efc"] == 40 and h["seed"] == 7
    sizes = [n, 4 * n, 4 * n * (1 + 2 * M), 4 * U * (1 + M), 4 * n * 40]
    at = ALIGN
    for s, (off, size, checksum) in enumerate(h["table"]):
        assert (off, size) == (at, sizes[s]), NAMES[s]
        assert checksum == o.xxh64(data[off:off + size]), NAMES[s]
        at = -(-(off + size) // ALIGN) * ALIGN
    assert len(data) == h["table"][4][0] + h["table"][4][1]
    assert h["header_sum"] == o.xxh64(data[:216]) and data[224:ALIGN] == bytes(ALIGN - 224)
    sec = [data[off:off + size] for off, size, _ in h["table"]]
    assert np.array_equal(np.frombuffer(sec[0], np.uint8), g["levels"])
    offsets = np.frombuffer(sec[1], np.uint32)
    assert np.array_equal(offsets, np.concatenate([[0], np.cumsum(g["levels"].astype(np.int64))[:-1]]))
    assert np.array_equal(np.frombuffer(sec[2], np.uint32), g["layer0"])
    assert np.array_equal(np.frombuffer(sec[3], np.uint32), g["upper"])
    rows = np.frombuffer(sec[4], np.float32).reshape(n, 40)
    assert np.array_equal(rows[:, :dim], g["vectors"]) and not rows[:, dim:].view(np.uint32).any()
    for s in range(4):  # gaps are zero
        end = h["table"][s][0] + h["table"][s][1]
        assert data[end:h["table"][s + 1][0]] == bytes(h["table"][s + 1][0] - end)


def test_round_trip_and_add_after_load(tmp_path):
    index = build()
    path = tmp_path / "g.vdx2"
    index.save(path, version=2)
    loaded = vdx.HnswIndex.load(path)
    assert_same(index, loaded)
    assert loaded.params() == index.params()
    extra = datasets.integer_matrix(60, 37, 9) + 1
    index.add(extra)
    loaded.add(extra)  # the level generator is restored: the same graph grows
    assert_same(index, loaded)
    # v1 of the same graph loads to the same sections
    index.save(tmp_path / "g.vdx", version=1)
    assert_same(index, vdx.HnswIndex.load(tmp_path / "g.vdx"))


def test_empty_index(tmp_path):
    index = vdx.HnswIndex(8, vdx.Metric.L2, M=4)
    index.save(tmp_path / "e.vdx2", version=2)
    assert (tmp_path / "e.vdx2").stat().st_size == ALIGN
    assert vdx.HnswIndex.load(tmp_path / "e.vdx2").size() == 0


def test_damage_is_detected(tmp_path):
    index = build(n=500)
    path = tmp_path / "g.vdx2"
    index.save(path, version=2)
    data = path.read_bytes()
    h = parse(data)
    bad = tmp_path / "bad.vdx2"
    for s, (off, size, _) in enumerate(h["table"]):
        if size == 0:
            continue
        d = bytearray(data)
        d[off + size // 2] ^= 0x10
        bad.write_bytes(bytes(d))
        with pytest.raises(RuntimeError, match=f"{NAMES[s]} section checksum mismatch|links|level|NaN|padding|upper block"):
            vdx.HnswIndex.load(bad)
    d = bytearray(data)
    d[20] ^= 1
    bad.write_bytes(bytes(d))
    with pytest.raises(RuntimeError, match="header checksum mismatch"):
        vdx.HnswIndex.load(bad)
    bad.write_bytes(data[:-4])
    with pytest.raises(RuntimeError, match="truncated or corrupted") as err:
        vdx.HnswIndex.load(bad)
    assert str(err.value).startswith(str(bad))
    bad.write_bytes(data[:100])
    with pytest.raises(RuntimeError, match="too small"):
        vdx.HnswIndex.load(bad)


def test_hostile_contents_are_rejected_after_resealing(tmp_path):
    index = build(n=500)
    path = tmp_path / "g.vdx2"
    index.save(path, version=2)
    data = path.read_bytes()
    h = parse(data)
    g = sections(index)
    M = g["M"]
    b0 = 1 + 2 * M
    (lv_off, _, _), (of_off, _, _), (l0_off, _, _), _, (vec_off, _, _) = h["table"]
    bad = tmp_path / "bad.vdx2"

    def rejected(patch, message):
        d = bytearray(data)
        patch(d)
        bad.write_bytes(reseal(d))
        with pytest.raises(RuntimeError, match=message):
            vdx.HnswIndex.load(bad)

    node = int(np.nonzero(g["layer0"][::b0] >= 2)[0][0])
    blk = l0_off + 4 * node * b0
    rejected(lambda d: d.__setitem__(slice(blk + 8, blk + 12), d[blk + 4:blk + 8]),
             f"node {node} links to node \\d+ twice on layer 0")
    rejected(lambda d: struct.pack_into("<I", d, blk + 4, 500), "but there are only 500 nodes")
    rejected(lambda d: struct.pack_into("<I", d, blk, 2 * M + 1), f"has {2 * M + 1} links on layer 0")
    rejected(lambda d: struct.pack_into("<f", d, vec_off + 4, float("nan")), "NaN or infinity")
    rejected(lambda d: struct.pack_into("<f", d, vec_off + 4 * 37, 1.0), "nonzero padding component")
    rejected(lambda d: struct.pack_into("<I", d, of_off + 4 * 499, 1), "first upper block is recorded")
    low = int(np.nonzero(g["levels"] == 0)[0][-1])
    rejected(lambda d: d.__setitem__(lv_off + low, 200), "above the max level")
    rejected(lambda d: struct.pack_into("<Q", d, 80, 44), "row length 44")
    rejected(lambda d: struct.pack_into("<Q", d, 96 + 24 * 2, l0_off + ALIGN), "layer-0 links section is at")
    rejected(lambda d: struct.pack_into("<I", d, 8, 3), "format version 3 is not supported")
    rejected(lambda d: d.__setitem__(1000, 1), "reserved bytes are not zero")


def test_save_rejects_unknown_versions(tmp_path):
    with pytest.raises(ValueError, match="format version 3"):
        build(n=20).save(tmp_path / "x", version=3)
