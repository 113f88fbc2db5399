"""GPU-assisted build (vdx_gpu_hnsw_build_add): the batched add with the GPU as the layer-0
searcher. The GPU's insertion searches equal the oracle's restatement of vecdex's search, so on
one thread the build equals the batched build driven by the oracle node for node; a batch of one
builds vecdex's serial graph (golden fingerprints, tests/hnsw_parallel_test.cpp:146-169)."""

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import _native as nat
from hnsw_vector_search_amd import datasets
from oracle import oracle as o
from batch_searchers import OracleSearcher, add_batched, sections

pytestmark = pytest.mark.gpu

GREEDY_FP = 6133164406513967666   # tests/hnsw_parallel_test.cpp:164
TOP10_FP = 15666754926265556722   # tests/hnsw_parallel_test.cpp:165

GRAPH_KEYS = ("levels", "layer0", "upper", "vectors", "max_level", "entry_point", "upper_blocks")


def fingerprint(out, cnt):
    return o.fingerprint([(out[i]["id"][:cnt[i]], out[i]["distance"][:cnt[i]]) for i in range(len(cnt))])


def assert_same_graph(a, b):
    a, b = sections(a), sections(b)
    for key in GRAPH_KEYS:
        assert np.array_equal(a[key], b[key]), key


def test_batch_of_one_on_the_gpu_builds_the_serial_graph():
    print("case: GPU build, batch of one, 1000 x 16", flush=True)
    base = datasets.integer_matrix(1000, 16, 1)
    queries = datasets.integer_matrix(200, 16, 2)
    serial = vdx.HnswIndex(16, vdx.Metric.L2, M=8, ef_construction=64, seed=42)
    serial.add(base)
    gpu = vdx.HnswIndex(16, vdx.Metric.L2, M=8, ef_construction=64, seed=42)
    gpu.add(base, 1, gpu_assisted=True, plan=nat.BuildPlan(0, 1, 1, 0.0))
    assert_same_graph(serial, gpu)
    # the device copy kept current batch by batch serves the golden fingerprints
    assert fingerprint(*gpu.search_arrays(queries, 1, 1)) == GREEDY_FP
    assert fingerprint(*gpu.search_arrays(queries, 10, 10)) == TOP10_FP


@pytest.mark.parametrize("dim,metric,M", [(128, vdx.Metric.L2, 12), (128, vdx.Metric.Cosine, 12),
                                          (100, vdx.Metric.Cosine, 32), (37, vdx.Metric.InnerProduct, 8)])
def test_gpu_batches_equal_oracle_batches(dim, metric, M):
    """The GPU's searches and neighbour choices against the oracle's searches and the host's
    choices: on one thread the two batched builds must agree node for node."""
    print(f"case: GPU build vs oracle-driven batched build, {dim}-d metric {metric} M={M}, 12000 rows", flush=True)
    if dim == 128:
        base = datasets.gmm64().sample(12_000, 1) + 1
    else:
        base = datasets.AnisotropicGaussian(dim, 0.8, seed=4).sample(12_000, 1)
    plan = nat.BuildPlan(1000, 128, 2048, 0.05)
    ref = vdx.HnswIndex(dim, metric, M=M, ef_construction=100, seed=3)
    assert add_batched(ref, base, 1, plan, OracleSearcher()) == 0
    gpu = vdx.HnswIndex(dim, metric, M=M, ef_construction=100, seed=3)
    gpu.add(base, 1, gpu_assisted=True, plan=plan)
    assert_same_graph(ref, gpu)


def test_threaded_gpu_build_recall_and_device_copy():
    print("case: threaded GPU build, 60000 x 128, M=16, efC=200", flush=True)
    mix = datasets.gmm1024()
    base, queries = mix.sample(60_000, 1), mix.sample(300, 2)
    host = vdx.HnswIndex(128, vdx.Metric.L2, M=16, ef_construction=200, seed=0)
    host.add(base, num_threads=8)
    gpu = vdx.HnswIndex(128, vdx.Metric.L2, M=16, ef_construction=200, seed=0)
    gpu.add(base[:30_000], num_threads=8, gpu_assisted=True)
    gpu.add(base[30_000:], num_threads=8, gpu_assisted=True)  # extends a GPU-built graph
    report = gpu.inspect()
    assert report.violations == [] and report.reachable >= host.inspect().reachable - 60
    truth, _ = o.bruteforce_f64(base, queries, 10)
    for ef in (16, 32, 64):
        r_host = o.recall_at_k(host.search_arrays(queries, 10, ef)[0]["id"], truth, 10)
        got, cnt = gpu.search_arrays(queries, 10, ef)
        r_gpu = o.recall_at_k(got["id"], truth, 10)
        assert r_gpu >= r_host - 0.01, (ef, r_host, r_gpu)
        # the incrementally synced device copy is exact: results equal the oracle's on the graph
        ref, ref_cnt = o.hnsw_search(gpu.graph_arrays(), queries, 10, ef)
        np.testing.assert_array_equal(cnt, ref_cnt)
        assert got.tobytes() == ref.tobytes()


def test_build_rejects_what_the_gpu_path_cannot_search():
    print("case: GPU build limits", flush=True)
    index = vdx.HnswIndex(16, vdx.Metric.L2, M=8, ef_construction=9000, seed=1)
    with pytest.raises(ValueError, match="ef_construction 9000 is above"):
        index.add(datasets.integer_matrix(50, 16, 1), gpu_assisted=True, plan=nat.BuildPlan(0, 8, 8, 0.0))
