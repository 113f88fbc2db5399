"""The oracle's expansion budget (vdxo_hnsw_search_batch_bounded; no vecdex counterpart).

Pinned three ways: a budget no query reaches is vecdex's search (the golden fingerprints of
tests/hnsw_parallel_test.cpp:164-165 hold through the bounded entry point); each query takes
min(budget, its own expansion count) expansions; and an independent restatement in Python
(heapq, one distance call per pair, vecdex's greedy descent and search_layer of
src/hnsw.cpp:416-482 with the stop after B expansions) gives the same ids and distance bits
for every budget from 0 up on a small graph."""

import heapq

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import datasets
from oracle import oracle as o

GREEDY_FP = 6133164406513967666   # tests/hnsw_parallel_test.cpp:164
TOP10_FP = 15666754926265556722   # tests/hnsw_parallel_test.cpp:165


def fp(out, cnt):
    return o.fingerprint([(out[i]["id"][:cnt[i]], out[i]["distance"][:cnt[i]]) for i in range(len(out))])


@pytest.fixture(scope="module")
def golden():
    base = datasets.integer_matrix(1000, 16, 1)
    queries = datasets.integer_matrix(200, 16, 2)
    index = vdx.HnswIndex(16, vdx.Metric.L2, M=8, ef_construction=64, seed=42)
    index.add(base, num_threads=1)
    return index, queries  # graph_arrays() are views into the index: keep it alive


def test_an_unreached_budget_is_vecdex_search(golden):
    index, queries = golden
    g = index.graph_arrays()
    for k, ef, want in ((1, 1, GREEDY_FP), (10, 10, TOP10_FP)):
        out, cnt, taken = o.hnsw_search(g, queries, k, ef, threads=4, max_expansions=0)
        assert fp(out, cnt) == want
        big = int(taken.max())
        out2, cnt2, taken2 = o.hnsw_search(g, queries, k, ef, threads=4, max_expansions=big)
        assert fp(out2, cnt2) == want and (taken2 == taken).all()
        assert taken.min() >= 1


def test_expansions_taken_are_the_budget_or_fewer(golden):
    index, queries = golden
    g = index.graph_arrays()
    _, _, natural = o.hnsw_search(g, queries, 10, 40, threads=4, max_expansions=0)
    for budget in (1, 5, 17, int(np.median(natural)), int(natural.max()) - 1):
        out, cnt, taken = o.hnsw_search(g, queries, 10, 40, threads=4, max_expansions=budget)
        np.testing.assert_array_equal(taken, np.minimum(natural, budget))
        full, full_cnt = o.hnsw_search(g, queries, 10, 40, threads=4)
        same = taken == natural
        assert (out[same].tobytes() == full[same].tobytes()) and (cnt[same] == full_cnt[same]).all()


def restated(g, query, k, ef, budget):
    """search_unchecked with the budget, written independently of the C oracle."""
    dim, M = g["dim"], g["M"]
    vec, layer0, upper = g["vectors"], g["layer0"], g["upper"]
    offsets = np.concatenate([[0], np.cumsum(g["levels"].astype(np.int64))])
    dist = lambda i: np.float32(o.distance("l2_avx2order", query, vec[i]))

    def links(node, layer):
        if layer == 0:
            b = layer0[node * (1 + 2 * M):(node + 1) * (1 + 2 * M)]
        else:
            at = (offsets[node] + layer - 1) * (1 + M)
            b = upper[at:at + 1 + M]
        return [int(x) for x in b[1:1 + b[0]]]

    cur, cur_d = int(g["entry_point"]), dist(int(g["entry_point"]))
    for layer in range(int(g["max_level"]), 0, -1):
        moved = True
        while moved:
            moved = False
            for nb in links(cur, layer):
                d = dist(nb)
                if d < cur_d:
                    cur, cur_d, moved = nb, d, True
    efp = max(ef, k)
    seen = {cur}
    cand = [(cur_d, cur)]                # min-heap by (distance, id): `closer`
    res = [(-cur_d, -cur)]               # max-heap of the kept results
    taken = 0
    while cand:
        d, node = heapq.heappop(cand)
        worst = (-res[0][0], -res[0][1])
        if len(res) >= efp and worst < (d, node):
            break
        if taken >= budget:
            break
        taken += 1
        for nb in links(node, 0):
            if nb in seen:
                continue
            seen.add(nb)
            dn = dist(nb)
            worst = (-res[0][0], -res[0][1])
            if len(res) < efp or (dn, nb) < worst:
                heapq.heappush(cand, (dn, nb))
                heapq.heappush(res, (-dn, -nb))
                if len(res) > efp:
                    heapq.heappop(res)
    found = sorted((-a, -b) for a, b in res)[:k]
    return [(np.float32(d), i) for d, i in found], taken


def test_matches_an_independent_restatement_for_every_budget():
    rng = np.random.default_rng(3)
    base = rng.uniform(-1, 1, (400, 8)).astype(np.float32)
    queries = rng.uniform(-1, 1, (6, 8)).astype(np.float32)
    index = vdx.HnswIndex(8, vdx.Metric.L2, M=6, ef_construction=40, seed=5)
    index.add(base, num_threads=1)
    g = index.graph_arrays()
    assert g["max_level"] >= 1
    k, ef = 5, 12
    _, _, natural = o.hnsw_search(g, queries, k, ef, threads=1, max_expansions=0)
    for budget in range(0, int(natural.max()) + 2):
        out, cnt, taken = o.hnsw_search(g, queries, k, ef, threads=1, max_expansions=budget or 10**9)
        for q in range(len(queries)):
            want, want_taken = restated(g, queries[q], k, ef, budget if budget else 10**9)
            got = [(out[q]["distance"][i], int(out[q]["id"][i])) for i in range(cnt[q])]
            assert [i for _, i in got] == [i for _, i in want], (budget, q)
            assert np.array([d for d, _ in got], np.float32).tobytes() == \
                np.array([d for d, _ in want], np.float32).tobytes()
            assert taken[q] == want_taken
