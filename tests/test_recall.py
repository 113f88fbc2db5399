"""recall_at_k / threshold_recall_at_k through the C ABI, pinned to the known-answer cases of
vecdex's tests/recall_test.cpp:28-101 and cross-checked against the oracle's numpy restatement
on random result matrices. CPU only."""

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import _native as nat
from oracle import oracle as o


def results(ids):
    """Results with the given ids, distances 0, 1, 2, ... (recall_test.cpp:18-26)."""
    return [vdx.Neighbor(float(i), int(n)) for i, n in enumerate(ids)]


def at_distances(dists):
    """Results with the given distances, ids 0, 1, 2, ... (recall_test.cpp:58-66)."""
    return [vdx.Neighbor(float(np.float32(d)), i) for i, d in enumerate(dists)]


def test_recall_perfect_partial_and_zero():
    truth = [4, 8, 15, 16]
    assert vdx.recall_at_k(results([4, 8, 15, 16]), truth, 4) == 1.0
    assert vdx.recall_at_k(results([16, 15, 8, 4]), truth, 4) == 1.0
    assert vdx.recall_at_k(results([4, 99, 15, 98]), truth, 4) == 0.5
    assert vdx.recall_at_k(results([1, 2, 3, 5]), truth, 4) == 0.0


def test_recall_uses_only_the_first_k_of_each_list():
    assert vdx.recall_at_k(results([4, 15, 8]), [4, 8, 15, 16], 2) == 0.5


def test_recall_missing_results_count_as_misses():
    assert vdx.recall_at_k(results([2, 1]), [1, 2, 3, 4], 4) == 0.5
    assert vdx.recall_at_k([], [1, 2, 3, 4], 4) == 0.0


def test_recall_duplicate_results_count_once():
    assert vdx.recall_at_k(results([7, 7, 7]), [7, 8, 9], 3) == 1.0 / 3.0


def test_recall_rejects_bad_arguments():
    with pytest.raises(ValueError, match="k must be positive"):
        vdx.recall_at_k(results([1]), [1, 2], 0)
    with pytest.raises(ValueError, match="truth has 2 ids, fewer than k = 3"):
        vdx.recall_at_k(results([1]), [1, 2], 3)


def test_recall_negative_true_id_never_matches():
    # ids are compared by value (std::cmp_equal): -1 is not 0xFFFFFFFF
    assert vdx.recall_at_k([vdx.Neighbor(0.0, 0xFFFFFFFF)], [-1], 1) == 0.0


def test_threshold_counts_ties_with_the_kth_neighbor():
    assert vdx.threshold_recall_at_k(at_distances([1, 2, 5]), 5.0, 3) == 1.0
    assert vdx.threshold_recall_at_k(at_distances([1, 2, 6]), 5.0, 3) == 2.0 / 3.0
    assert vdx.threshold_recall_at_k(at_distances([7, 8, 9]), 5.0, 3) == 0.0


def test_threshold_first_k_and_missing_results():
    assert vdx.threshold_recall_at_k(at_distances([1, 9, 2]), 5.0, 2) == 0.5
    assert vdx.threshold_recall_at_k(at_distances([1]), 5.0, 4) == 0.25
    assert vdx.threshold_recall_at_k([], 5.0, 4) == 0.0


def test_threshold_tolerance_relative_above_one_absolute_below():
    assert vdx.threshold_recall_at_k(at_distances([100000.5]), 100000.0, 1) == 1.0
    assert vdx.threshold_recall_at_k(at_distances([100002.0]), 100000.0, 1) == 0.0
    assert vdx.threshold_recall_at_k(at_distances([5e-6]), 0.0, 1) == 1.0
    assert vdx.threshold_recall_at_k(at_distances([5e-5]), 0.0, 1) == 0.0
    assert vdx.threshold_recall_at_k(at_distances([100000.5]), 100000.0, 1, 0.0) == 0.0


@pytest.mark.parametrize("kth,k,tol", [(1.0, 0, 1e-5), (float("nan"), 1, 1e-5),
                                       (float("inf"), 1, 1e-5), (1.0, 1, -1e-5),
                                       (1.0, 1, float("nan"))])
def test_threshold_rejects_bad_arguments(kth, k, tol):
    with pytest.raises(ValueError, match="threshold_recall_at_k"):
        vdx.threshold_recall_at_k(at_distances([1]), kth, k, tol)


def test_means_agree_with_the_restatement():
    rng = np.random.default_rng(5)
    nq, stride, k = 300, 12, 10
    found = np.zeros((nq, stride), dtype=nat.NEIGHBOR_DTYPE)
    found["id"] = rng.integers(0, 40, size=(nq, stride))
    found["distance"] = np.sort(rng.integers(0, 50, size=(nq, stride)).astype(np.float32), axis=1)
    truth = np.stack([rng.permutation(40)[:k] for _ in range(nq)]).astype(np.int32)
    counts = rng.integers(0, stride + 1, size=nq).astype(np.uint32)
    kth = rng.integers(0, 50, size=nq).astype(np.float32)
    # oracle over rows trimmed to their counts
    want_ids = np.mean([o.recall_at_k(found["id"][q, :counts[q]][None, :], truth[q][None, :], k)
                        for q in range(nq)])
    want_thr = np.mean([o.threshold_recall_at_k(found["distance"][q, :counts[q]][None, :],
                                                kth[q:q + 1], k) for q in range(nq)])
    assert vdx.mean_recall_at_k(found, truth, k, counts) == pytest.approx(want_ids, abs=1e-12)
    assert vdx.mean_threshold_recall_at_k(found, kth, k, counts=counts) == pytest.approx(want_thr, abs=1e-12)
    # per-row entry points agree with the means
    rows = [vdx.recall_at_k(found[q, :counts[q]], truth[q], k) for q in range(nq)]
    assert np.mean(rows) == pytest.approx(want_ids, abs=1e-12)
