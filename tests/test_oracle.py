"""The oracle against the reference's known answers and against what the mathematics fixes.

Each check is chosen so that a plausible restatement mistake (a dropped term, a wrong reduction
order, a wrong sign or tie-break) fails it."""

import numpy as np
import pytest

from oracle import oracle as o
from hnsw_vector_search_amd import datasets

PRIME32 = 2654435761
PRIME64 = 11400714785074694797


def sanity_buffer(size):
    """XSUM_fillTestBuffer of xxHash (tests/xxhash64_test.cpp:20-29)."""
    out = bytearray(size)
    gen = PRIME32
    for i in range(size):
        out[i] = gen >> 56
        gen = (gen * PRIME64) & ((1 << 64) - 1)
    return bytes(out)


# tests/xxhash64_test.cpp:34-56 (xxHash v0.8.3 XSUM_XXH64_testdata)
XXH64_VECTORS = [
    (0, 0, 0xEF46DB3751D8E999), (0, PRIME32, 0xAC75FDA2929B17EF),
    (1, 0, 0xE934A84ADB052768), (1, PRIME32, 0x5014607643A9B4C3),
    (4, 0, 0x9136A0DCA57457EE), (14, 0, 0x8282DCC4994E35C8),
    (14, PRIME32, 0xC3BD6BF63DEB6DF0), (222, 0, 0xB641AE8CB691C174),
    (222, PRIME32, 0x20CB8AB7AE10C14A),
]


@pytest.mark.parametrize("length,seed,expected", XXH64_VECTORS)
def test_xxh64_reference_vectors(length, seed, expected):
    assert o.xxh64(sanity_buffer(222)[:length], seed) == expected


@pytest.mark.parametrize("length,low", [(4096, 0x85F4E70B), (100_003, 0x1682A665),
                                        (1_048_583, 0x4B6FAE5C)])
def test_xxh64_low_bits_match_zstd(length, low):
    # tests/xxhash64_test.cpp:61-75
    assert o.xxh64(sanity_buffer(length)) & 0xFFFFFFFF == low


def test_distance_known_values():
    # tests/distance_test.cpp:16-58
    for kind in ("avx2order", "scalar", "fast"):
        assert o.distance("l2_" + kind, [0, 0], [3, 4]) == 25.0
        assert o.distance("l2_" + kind, [3, 4], [3, 4]) == 0.0
        assert o.distance("l2_" + kind, [], []) == 0.0
        assert o.distance("ip_" + kind, [1, 2, 3], [4, -5, 6]) == 12.0
        a = np.arange(128, dtype=np.float32)
        assert o.distance("l2_" + kind, a, 127 - a) == float(((2 * np.arange(128) - 127) ** 2).sum())


def test_orders_agree_bitwise_on_small_integers():
    # tests/kernels_test.cpp:144-161: every summation order is exact on small integers
    rng = np.random.default_rng(3)
    for dim in list(range(0, 101)) + [128, 256, 384, 768]:
        a = rng.integers(-50, 50, dim).astype(np.float32)
        b = rng.integers(-50, 50, dim).astype(np.float32)
        exact_l2 = float(((a.astype(np.int64) - b) ** 2).sum())
        exact_ip = float((a.astype(np.int64) * b).sum())
        for kind in ("avx2order", "scalar", "fast"):
            assert o.distance("l2_" + kind, a, b) == exact_l2
            assert o.distance("ip_" + kind, a, b) == exact_ip


def test_avx2_restatement_matches_intrinsics_bitwise_on_floats():
    """The scalar restatement of the AVX2 order against the real AVX2+FMA instructions: any
    slip in the 32-slot accumulation or the reduction tree changes low bits here."""
    if not o.has_avx2():
        pytest.skip("no AVX2 on this host")
    rng = np.random.default_rng(4)
    for dim in list(range(0, 101)) + [127, 128, 129, 200, 256, 287]:
        for _ in range(5):
            a = rng.standard_normal(dim).astype(np.float32)
            b = rng.standard_normal(dim).astype(np.float32)
            assert o.distance("l2_avx2order", a, b) == o.distance("l2_fast", a, b), dim
            assert o.distance("ip_avx2order", a, b) == o.distance("ip_fast", a, b), dim


def test_avx2_order_differs_from_left_to_right_order():
    """The order matters: on float data it is not the scalar kernel's."""
    rng = np.random.default_rng(5)
    differs = 0
    for _ in range(200):
        a = rng.standard_normal(100).astype(np.float32)
        b = rng.standard_normal(100).astype(np.float32)
        differs += o.distance("ip_avx2order", a, b) != o.distance("ip_scalar", a, b)
    assert differs > 20


def test_distances_within_rounding_of_float64():
    # tests/kernels_test.cpp:78-81: |error| <= 2 (d + 2) eps * magnitude
    rng = np.random.default_rng(6)
    eps = np.finfo(np.float32).eps
    for dim in (1, 7, 33, 100, 128, 300):
        a = rng.uniform(-1, 1, dim).astype(np.float32)
        b = rng.uniform(-1, 1, dim).astype(np.float32)
        l2 = float(((a.astype(np.float64) - b) ** 2).sum())
        ip = float((a.astype(np.float64) * b).sum())
        mag = float((np.abs(a.astype(np.float64)) * np.abs(b)).sum())
        assert abs(o.distance("l2_avx2order", a, b) - l2) <= 2 * (dim + 2) * eps * l2 + 1e-30
        assert abs(o.distance("ip_avx2order", a, b) - ip) <= 2 * (dim + 2) * eps * mag


def test_normalize_gives_unit_length():
    v = o.normalize(np.array([[3, 4], [10, 0]], dtype=np.float32))
    np.testing.assert_array_equal(v, np.array([[3 / 5, 4 / 5], [1, 0]], dtype=np.float32))


def test_bruteforce_line_cases():
    # tests/brute_force_test.cpp:46-73
    base = np.array([[10], [0], [7], [3], [-2]], dtype=np.float32)
    out = o.bruteforce_f32(base, np.array([[2]], np.float32), 3)
    assert out["id"].tolist() == [[3, 1, 4]] and out["distance"].tolist() == [[1, 4, 16]]
    out = o.bruteforce_f32(np.array([[5], [1]], np.float32), np.array([[0]], np.float32), 10)
    assert out["id"].tolist() == [[1, 0]] and out["distance"].tolist() == [[1, 25]]
    ties = np.array([[4], [-4], [4], [9], [-4]], dtype=np.float32)
    out = o.bruteforce_f32(ties, np.array([[0]], np.float32), 4)
    assert out["id"].tolist() == [[0, 1, 2, 4]]
    ids, dist = o.bruteforce_f64(ties, np.array([[0]], np.float32), 2)
    assert ids.tolist() == [[0, 1]] and dist.tolist() == [[16, 16]]


def test_bruteforce_inner_product_and_cosine_cases():
    # tests/brute_force_test.cpp:158-175
    base = np.array([[1, 0], [0, 2], [3, 3]], dtype=np.float32)
    out = o.bruteforce_f32(base, np.array([[1, 1]], np.float32), 3, metric=1)
    assert out["id"].tolist() == [[2, 1, 0]] and out["distance"].tolist() == [[-5, -1, 0]]
    stored = o.normalize(np.array([[10, 0], [0, 0.5], [-3, 0]], dtype=np.float32))
    q = o.normalize(np.array([[2, 0]], dtype=np.float32))
    out = o.bruteforce_f32(stored, q, 3, metric=2)
    assert out["id"].tolist() == [[0, 1, 2]] and out["distance"].tolist() == [[0, 1, 2]]


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_bruteforce_matches_numpy_float64(metric):
    """Exact top-k: float32 brute force within rounding of an independent numpy float64 one
    (tests/brute_force_test.cpp:218-262)."""
    rng = np.random.default_rng(5)
    base = rng.uniform(-1, 1, (300, 24)).astype(np.float32)
    q = rng.uniform(-1, 1, (10, 24)).astype(np.float32)
    b64, q64 = base.astype(np.float64), q.astype(np.float64)
    if metric == 0:
        ref = ((q64[:, None, :] - b64[None]) ** 2).sum(-1)
    elif metric == 1:
        ref = 1 - q64 @ b64.T
    else:
        ref = 1 - (q64 @ b64.T) / np.outer(np.linalg.norm(q64, axis=1), np.linalg.norm(b64, axis=1))
    stored, queries = (o.normalize(base), o.normalize(q)) if metric == 2 else (base, q)
    ids64, d64 = o.bruteforce_f64(base, q, 50, metric)
    np.testing.assert_allclose(d64, np.sort(ref, axis=1)[:, :50], rtol=1e-12, atol=1e-12)
    out = o.bruteforce_f32(stored, queries, 50, metric)
    tol = lambda r: 1e-5 * (1 + np.abs(r))
    for r in range(10):
        kth = np.sort(ref[r])[49]
        assert np.all(ref[r][out["id"][r]] <= kth + tol(kth))
        np.testing.assert_allclose(out["distance"][r], ref[r][out["id"][r]], atol=1e-5, rtol=1e-5)


def test_recall_functions():
    # tests/recall_test.cpp: recall by id and by threshold
    assert o.recall_at_k([[1, 2, 3]], [[3, 2, 9]], 3) == pytest.approx(2 / 3)
    assert o.threshold_recall_at_k([[1.0, 2.0, 5.0]], [2.0], 3) == pytest.approx(2 / 3)


def test_fingerprint_mixes_ids_and_distance_bits():
    a = o.fingerprint([([1, 2], [1.0, 2.0])])
    assert a != o.fingerprint([([2, 1], [1.0, 2.0])])
    assert a != o.fingerprint([([1, 2], [1.0, np.nextafter(np.float32(2), np.float32(3))])])


def test_integer_matrix_matches_mt19937_top_bits():
    # python/tests/test_threads.py:24-29: NumPy's RandomState is mt19937
    m = datasets.integer_matrix(2, 3, 1)
    raw = np.random.RandomState(1).randint(0, 2**32, size=6, dtype=np.uint32)
    np.testing.assert_array_equal(m.ravel(), (raw >> 24).astype(np.int32) - 128)
    assert m.min() >= -128 and m.max() <= 127
