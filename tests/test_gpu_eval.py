"""The vecdex_eval-compatible sweep on the GPU path, over a small TEXMEX dataset written here:
its CSV rows, its recall against float64 ground truth, and its results fingerprints equal to
those of the oracle's CPU search over the same saved graph."""

import re

import numpy as np
import pytest

import hnsw_vector_search_amd as vdx
from hnsw_vector_search_amd import datasets
from hnsw_vector_search_amd import eval as ev
from hnsw_vector_search_amd.vecs import write_fvecs, write_ivecs
from oracle import oracle as o

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("device_resident", [False, True])
def test_eval_sweep(tmp_path, capsys, device_resident):
    print(f"case: eval sweep, device_resident={device_resident}", flush=True)
    base = datasets.integer_matrix(3000, 24, 1)
    queries = datasets.integer_matrix(80, 24, 2)
    truth = o.bruteforce_f64(base, queries, 10)[0].astype(np.int32)
    write_fvecs(tmp_path / "toy_base.fvecs", base)
    write_fvecs(tmp_path / "toy_query.fvecs", queries)
    write_ivecs(tmp_path / "toy_groundtruth.ivecs", truth)
    graph_path = tmp_path / "toy.vdx"
    args = [str(tmp_path), "toy", "--M", "8", "--ef-construction", "64", "--ef", "16,64",
            "--repeats", "3", "--save", str(graph_path)]
    if device_resident:
        args.append("--device-resident")
    assert ev.main(args) == 0
    out = capsys.readouterr()
    lines = out.out.strip().splitlines()
    lines = lines[next(i for i, ln in enumerate(lines) if ln.startswith("lib,")):]
    assert lines[0] == "lib,dataset,n,dim,M,ef_construction,build_s,memory_mb,k,ef,recall,recall_dist,qps,threads"
    rows = [dict(zip(lines[0].split(","), ln.split(","))) for ln in lines[1:]]
    assert [r["ef"] for r in rows] == ["16", "64"]
    assert all(r["lib"] == "vecdex-mi355x" and r["n"] == "3000" and r["dim"] == "24" for r in rows)
    assert float(rows[1]["recall"]) >= 0.95 and float(rows[1]["recall_dist"]) >= float(rows[1]["recall"])
    fps = dict(re.findall(r"ef (\d+): results fingerprint ([0-9a-f]{16})", out.err))
    saved = vdx.HnswIndex.load(graph_path)  # graph_arrays() are views into it: keep it alive
    g = saved.graph_arrays()
    for ef in (16, 64):
        ref, cnt = o.hnsw_search(g, queries, 10, ef)
        assert fps[str(ef)] == f"{ev.fingerprint(ref, cnt):016x}"
