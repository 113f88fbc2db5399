"""The oracle-stop ceiling of learned termination (learned.stop_ceiling), pinned on the CPU
against brute force: over every choice of per-query stop (the first expansion reaching h hits,
any h up to the query's final hits), the fewest total expansions whose mean recall reaches the
target; and the label event (first expansion with the final hits) on hand-made traces."""

import itertools

import numpy as np

from hnsw_vector_search_amd import learned


def exhaustive(hits, steps, final, target, k=10):
    nq, cap = hits.shape

    def first(q, h):
        return 0 if h == 0 else next((s + 1 for s in range(min(steps[q], cap)) if hits[q, s] >= h), steps[q])

    need = int(np.ceil(target * k * nq - 1e-9))
    return min(sum(first(q, h) for q, h in enumerate(hs))
               for hs in itertools.product(*[range(f + 1) for f in final]) if sum(hs) >= need) / nq


def test_hand_made_traces():
    hits = np.array([[1, 3, 5, 8, 9, 10], [0, 2, 9, 9, 9, 9], [4, 4, 4, 4, 4, 4]], np.uint8)
    c = learned.stop_ceiling(hits, np.array([6, 6, 6]), np.array([10, 9, 4]), target=23 / 30)
    assert c["fixed_expansions"] == 6.0
    assert c["oracle_stop_expansions"] == round((6 + 3 + 1) / 3, 2)  # first expansion with the final hits
    assert c["budgeted_expansions"] == round(10 / 3, 2)
    # a target above the fixed recall cannot be met by stopping early
    assert learned.stop_ceiling(hits, np.array([6, 6, 6]), np.array([10, 9, 4]), 0.9)["budgeted_expansions"] is None


def test_budgeted_stop_is_the_exhaustive_optimum():
    rng = np.random.default_rng(1)
    for trial in range(150):
        nq, cap = 3, 8
        hits = np.minimum(np.cumsum(rng.integers(0, 3, (nq, cap)), axis=1), 10).astype(np.uint8)
        drop = rng.random((nq, cap)) < 0.1
        hits[drop] = np.maximum(hits[drop].astype(int) - 1, 0)  # hits may fall back for a step
        steps = rng.integers(5, cap + 3, nq)  # searches longer than the recorded trace too
        final = hits[np.arange(nq), np.minimum(steps, cap) - 1].astype(np.int64)
        target = rng.integers(max(0, final.sum() - 7), final.sum() + 1) / (10 * nq)
        got = learned.stop_ceiling(hits, steps, final, target)
        assert got["budgeted_expansions"] == round(exhaustive(hits, steps, final, target), 2), trial
        assert got["budgeted_expansions"] <= got["oracle_stop_expansions"] + 1e-9 or target > final.mean() / 10


def test_per_query_stops_realise_the_ceiling():
    """The per-query stops the DP returns: their mean is the reported ceiling, and the hits each
    query holds at its stop reach the target in total."""
    rng = np.random.default_rng(2)
    for trial in range(60):
        nq, cap = int(rng.integers(2, 6)), 10
        hits = np.minimum(np.cumsum(rng.integers(0, 3, (nq, cap)), axis=1), 10).astype(np.uint8)
        steps = rng.integers(4, cap + 1, nq)
        final = hits[np.arange(nq), steps - 1].astype(np.int64)
        target = rng.integers(max(0, final.sum() - 9), final.sum() + 1) / (10 * nq)
        got = learned.stop_ceiling(hits, steps, final, target, per_query=True)
        stops = got["budgeted_stops"].astype(np.int64)
        assert round(stops.mean(), 2) == got["budgeted_expansions"], trial
        held = [0 if s == 0 else int(hits[q, s - 1]) for q, s in enumerate(stops)]
        assert sum(held) >= int(np.ceil(target * 10 * nq - 1e-9)), trial
        assert (got["oracle_stops"].astype(np.int64) <= steps).all()
        assert round(float(got["oracle_stops"].mean()), 2) == got["oracle_stop_expansions"]
