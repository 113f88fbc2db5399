"""FORMAT v2 streamed straight to HBM (vdx_gpu_hnsw_load, HnswDeviceIndex): no host graph, each
section copied through pinned chunks after its checks. Results must equal the host graph's GPU
search (itself equal to the oracle's) on the same graph, bit for bit."""

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import _native as nat
from hnsw_vector_search_amd import datasets
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def identical(a, b, queries, k, ef):
    ra, ca = a.search_arrays(queries, k, ef)
    rb, cb = b.search_arrays(queries, k, ef)
    np.testing.assert_array_equal(ca, cb)
    assert ra.tobytes() == rb.tobytes(), (k, ef)


@pytest.mark.parametrize("dim,metric", [(128, vdx.Metric.L2), (100, vdx.Metric.Cosine), (37, vdx.Metric.InnerProduct)])
def test_device_load_serves_the_saved_graph(tmp_path, dim, metric):
    print(f"case: v2 device load, 8000 x {dim}, metric {metric}", flush=True)
    gen = datasets.AnisotropicGaussian(dim, 0.8, seed=1)
    index = vdx.HnswIndex(dim, metric, M=12, ef_construction=80, seed=2)
    index.add(gen.sample(8000, 1), num_threads=8)
    path = tmp_path / "g.vdx2"
    index.save(path, version=2)
    dev = vdx.HnswDeviceIndex.load(path)
    assert (dev.dim(), dev.size(), dev.metric()) == (dim, 8000, metric)
    queries = gen.sample(200, 2)
    for k, ef in ((1, 1), (10, 32), (10, 128), (50, 600)):
        identical(index, dev, queries, k, ef)
    prepared = o.normalize(queries) if metric == vdx.Metric.Cosine else queries
    ref, ref_cnt = o.hnsw_search(index.graph_arrays(), prepared, 10, 64)
    got, cnt = dev.search_arrays(queries, 10, 64)
    np.testing.assert_array_equal(cnt, ref_cnt)
    assert got.tobytes() == ref.tobytes()


def test_multi_chunk_load_and_batches(tmp_path):
    print("case: v2 device load across pinned chunks, 150000 x 128", flush=True)
    mix = datasets.gmm1024()
    index = vdx.HnswIndex(128, vdx.Metric.L2, M=16, ef_construction=100, seed=0)
    index.add(mix.sample(150_000, 1), num_threads=16)
    path = tmp_path / "big.vdx2"
    index.save(path, version=2)  # 77 MB of vectors: two 64 MB chunks
    dev = vdx.HnswDeviceIndex.load(path)
    queries = mix.sample(500, 2)
    identical(index, dev, queries, 10, 48)
    ctx = nat.Context.get(0)
    dq = nat.DeviceArray.from_numpy(queries, ctx)
    out = nat.DeviceArray((500, 10), nat.NEIGHBOR_DTYPE, ctx)
    cnt = nat.DeviceArray(500, np.uint32, ctx)
    batch = vdx.SearchBatch(dev, 500, 10, 64)
    batch.run(dq, 500, out, cnt, k=10, ef=48)
    nat.check(nat.lib.vdx_gpu_synchronize(ctx.handle))
    ref, _ = index.search_arrays(queries, 10, 48)
    assert out.numpy().tobytes() == ref.tobytes()


def test_damaged_file_is_refused(tmp_path):
    print("case: v2 device load of a damaged file", flush=True)
    index = vdx.HnswIndex(16, vdx.Metric.L2, M=8, ef_construction=40, seed=1)
    index.add(datasets.integer_matrix(3000, 16, 1))
    path = tmp_path / "g.vdx2"
    index.save(path, version=2)
    data = bytearray(path.read_bytes())
    data[-100] ^= 0x04  # a vector component: its section checksum no longer matches
    bad = tmp_path / "bad.vdx2"
    bad.write_bytes(bytes(data))
    with pytest.raises(RuntimeError, match="vectors section checksum mismatch") as err:
        vdx.HnswDeviceIndex.load(bad)
    assert str(err.value).startswith(str(bad))
    with pytest.raises(RuntimeError, match="format version 1: this loader reads version 2 files"):
        index.save(tmp_path / "g.vdx", version=1)
        vdx.HnswDeviceIndex.load(tmp_path / "g.vdx")
    big_m = vdx.HnswIndex(8, vdx.Metric.L2, M=40, ef_construction=40, seed=1)
    big_m.add(datasets.integer_matrix(100, 8, 1))
    big_m.save(tmp_path / "m.vdx2", version=2)
    with pytest.raises(ValueError, match="outside the GPU path's range"):
        vdx.HnswDeviceIndex.load(tmp_path / "m.vdx2")
