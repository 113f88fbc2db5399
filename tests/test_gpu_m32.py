"""M = 32 graphs on the 4,096-component mixture (the generator of the 10,000,000 x 128 base,
graph M=32, efConstruction=200), at 60,000 rows, against the oracle's restatement of vecdex's
search on the same graph.

At M = 32 a layer-0 block holds up to 2M = 64 links (docs/FORMAT.md:24-30): a full block
fills the beam's fresh-id and fresh-distance buffers (64 entries each) with every lane of the
wave, the link-block prefetch reads all 64 lanes, and the 8-bit and FP32 row loops run two
rounds of 32 candidates per expansion. Every (ef, k) below must give identical results
(ids and distance bits) over both row formats, in one launch and in uneven batches, and on a
GPU-assisted build of the same shape."""

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import _native as nat
from hnsw_vector_search_amd import datasets
from oracle import oracle as o

pytestmark = pytest.mark.gpu

N, NQ = 60_000, 300


def f32_rows():
    opts = nat.SearchOpts()
    opts.row_format = 1
    return opts


@pytest.fixture(scope="module")
def m32():
    print(f"case: building {N} x 128, M=32, efC=200 (4,096-component mixture)", flush=True)
    mix = datasets.gmm1024(4096)
    base, queries = mix.sample(N, 1), mix.sample(NQ, 2)
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=32, ef_construction=200, seed=0)
    index.add(base, num_threads=16)
    return index, base, queries


def assert_matches_oracle(index, queries, k, ef, **kw):
    ref, ref_cnt = o.hnsw_search(index.graph_arrays(), queries, k, ef, threads=16)
    got, cnt = index.search_arrays(queries, k, ef, **kw)
    np.testing.assert_array_equal(cnt, ref_cnt)
    for q in range(len(queries)):
        c = int(cnt[q])
        assert got[q, :c].tobytes() == ref[q, :c].tobytes(), f"query {q}, k {k}, ef {ef}"
    return got


def test_the_graph_has_full_layer0_blocks(m32):
    index, _, _ = m32
    g = index.graph_arrays()
    counts = g["layer0"].reshape(g["n"], -1)[:, 0]
    assert counts.max() == 64 and (counts == 64).sum() > 100
    assert index.gpu_u8_rows()  # integer data: the plain search reads the 8-bit rows


@pytest.mark.parametrize("ef", [16, 64, 512, 700])
@pytest.mark.parametrize("k", [1, 10, 100])
def test_identical_to_cpu_search_over_both_row_formats(m32, k, ef):
    print(f"case: M=32 k {k} ef {ef}", flush=True)
    index, _, queries = m32
    got = assert_matches_oracle(index, queries, k, ef)
    f32, _ = index.search_arrays(queries, k, ef, opts=f32_rows())
    assert f32.tobytes() == got.tobytes()


def test_recall_against_brute_force(m32):
    index, base, queries = m32
    truth = o.bruteforce_f32(base, queries, 10, threads=16)["id"]
    got, _ = index.search_arrays(queries, 10, 64)
    assert o.recall_at_k(got["id"], truth, 10) >= 0.95


def test_uneven_batch_equals_standalone_searches(m32):
    print("case: M=32 uneven batch", flush=True)
    index, _, queries = m32
    rng = np.random.default_rng(11)
    efs = np.array([16, 64, 256, 512, 700], dtype=np.uint32)[rng.integers(0, 5, NQ)]
    ks = np.array([1, 10, 100], dtype=np.uint32)[rng.integers(0, 3, NQ)]
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(queries, ctx)
    dks = nat.DeviceArray.from_numpy(ks, ctx)
    defs = nat.DeviceArray.from_numpy(efs, ctx)
    out = nat.DeviceArray((NQ, 100), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(NQ, np.uint32, ctx)
    batch = vdx.SearchBatch(index, NQ, 100, 700)
    batch.run(dq, NQ, out, cnt, ks=dks, efs=defs, graph=False)
    nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
    got, gcnt = out.numpy(), cnt.numpy()
    g = index.graph_arrays()
    for ef in np.unique(efs):
        for k in np.unique(ks):
            sel = np.nonzero((efs == ef) & (ks == k))[0]
            if len(sel) == 0:
                continue
            ref, ref_cnt = o.hnsw_search(g, queries[sel], int(k), int(ef), threads=16)
            np.testing.assert_array_equal(gcnt[sel], ref_cnt)
            for i, q in enumerate(sel):
                c = int(ref_cnt[i])
                assert got[q, :c].tobytes() == ref[i, :c].tobytes(), (q, ef, k)
    del batch


def test_gpu_assisted_build_at_m32(m32):
    """The same shape built GPU-assisted: a valid graph with the host build's recall, whose
    device copy (kept current batch by batch) serves the oracle's results."""
    print("case: M=32 GPU-assisted build", flush=True)
    host, base, queries = m32
    gpu = vdx.HnswIndex(128, vdx.Metric.L2, M=32, ef_construction=200, seed=0)
    gpu.add(base, num_threads=16, gpu_assisted=True)
    report = gpu.inspect()
    assert report.violations == [] and report.reachable == N
    for k, ef in ((10, 64), (100, 512)):
        assert_matches_oracle(gpu, queries, k, ef)
    truth = o.bruteforce_f32(base, queries, 10, threads=16)["id"]
    r_gpu = o.recall_at_k(gpu.search_arrays(queries, 10, 64)[0]["id"], truth, 10)
    r_host = o.recall_at_k(host.search_arrays(queries, 10, 64)[0]["id"], truth, 10)
    print(f"  recall@10 at ef 64: GPU-built {r_gpu:.4f}, host-built {r_host:.4f}", flush=True)
    assert r_gpu >= r_host - 0.01
