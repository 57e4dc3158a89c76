"""The C++ host program (cli/gvs_search.cpp over include/gvs_hnswlib.hpp):
argument and file errors on the CPU; on the GPU, its results equal the
library's search through the Python binding and its recall report matches
a recall computed here."""
import os
import subprocess

import numpy as np
import pytest

import gvs
from oracle import oracle as orc

CLI = os.path.join(os.path.dirname(gvs.LIB_PATH), "gvs_search")


def write_vecs(path, a):
    a = np.ascontiguousarray(a)
    with open(path, "wb") as f:
        for row in a:
            f.write(np.int32(row.shape[0]).tobytes())
            f.write(row.tobytes())


def test_cli_usage_and_missing_files(tmp_path):
    r = subprocess.run([CLI], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([CLI, str(tmp_path / "none.bin"), "16", "l2", str(tmp_path / "none.fvecs"), "10", "64"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "cannot read" in r.stderr


@pytest.mark.gpu
def test_cli_matches_library_and_reports_recall(tmp_path):
    x = gvs.gen_mixture(5000, 64, 16, 81, 82)
    q = gvs.gen_mixture(300, 64, 16, 81, 83)
    g = gvs.HostGraph.build_hnsw(x, M=12, ef_construction=100, nthreads=8)
    idx = str(tmp_path / "g.bin")
    g.save_hnswlib(idx)
    write_vecs(tmp_path / "q.fvecs", q)
    gt, _ = orc.bruteforce_f64(x, q, 10)
    write_vecs(tmp_path / "gt.ivecs", gt.astype(np.int32))
    print("case cli k=10 ef=48 mc=2", flush=True)
    r = subprocess.run([CLI, idx, "64", "l2", str(tmp_path / "q.fvecs"), "10", "48", "--mc", "2",
                        "--gt", str(tmp_path / "gt.ivecs"), "--out", str(tmp_path / "res")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ids = np.fromfile(tmp_path / "res.ids", dtype=np.int64).reshape(300, 10)
    ds = np.fromfile(tmp_path / "res.dists", dtype=np.float32).reshape(300, 10)
    dev = gvs.DeviceIndex(gvs.HostGraph.load_hnswlib(idx, 64))
    lab, dist = dev.search(q, k=10, ef=48, mc=2)
    np.testing.assert_array_equal(ids, lab)
    np.testing.assert_array_equal(ds, dist)
    rep = dict(line.split("=", 1) for line in r.stdout.split() if "=" in line)
    assert abs(float(rep["Recall@10"]) - orc.recall_at(lab, gt, 10)) < 1e-8
    assert abs(float(rep["Recall@1"]) - float(np.mean(lab[:, 0] == gt[:, 0]))) < 1e-8
