"""Write the golden vectors under tests/golden/ from the reference's vendored
hnswlib (oracle/_ref/ref_hnsw, compiled from the reference's own headers by
oracle/Makefile `ref`).  Run in a container that has the reference tree:

    make -C oracle ref && python tests/golden/make_golden.py

For each case the fixture holds the inputs (base vectors, queries), the
reference-built index in hnswlib's .bin format, and the reference's
searchKnnCloserFirst output (ids, dists) for several ef.  The reference is
built with -DNO_MANUAL_VECTORIZATION (scalar L2Sqr / InnerProduct).
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_hnsw")

CASES = [
    # name, n, dim, metric, M, efc, centres, normalize, nq, k, efs
    ("l2_n1000_d16_M8", 1000, 16, "l2", 8, 64, 12, False, 40, 10, (10, 40, 100)),
    ("ip_n1000_d24_M6", 1000, 24, "ip", 6, 48, 8, True, 40, 10, (10, 40, 100)),
    ("l2_n400_d8_M4", 400, 8, "l2", 4, 16, 4, False, 30, 5, (5, 20, 400)),
]


def mixture(n, dim, centres, seed_c, seed_p, normalize):
    rc = np.random.default_rng(seed_c)
    cen = rc.normal(0.0, 4.0, size=(centres, dim))
    rp = np.random.default_rng(seed_p)
    u = rp.integers(0, centres, size=n)
    x = (cen[u] + rp.normal(0.0, 1.0, size=(n, dim))).astype(np.float32)
    if normalize:
        x /= np.linalg.norm(x, axis=1, keepdims=True).astype(np.float32)
    return np.ascontiguousarray(x, dtype=np.float32)


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run `make -C oracle ref` where the reference tree exists")
    for name, n, dim, metric, M, efc, centres, norm, nq, k, efs in CASES:
        x = mixture(n, dim, centres, 11, 12, norm)
        q = mixture(nq, dim, centres, 11, 13, norm)
        with tempfile.TemporaryDirectory() as td:
            xf, qf = os.path.join(td, "x.f32"), os.path.join(td, "q.f32")
            x.tofile(xf)
            q.tofile(qf)
            idx = os.path.join(HERE, name + ".hnsw.bin")
            subprocess.run([REF, "build", xf, str(n), str(dim), metric, str(M), str(efc), "100", idx], check=True)
            out = {"x": x, "q": q, "k": np.int32(k), "M": np.int32(M), "efc": np.int32(efc),
                   "metric": np.int32(0 if metric == "l2" else 1), "efs": np.array(efs, dtype=np.int32)}
            for ef in efs:
                fi, fd = os.path.join(td, "i"), os.path.join(td, "d")
                subprocess.run([REF, "search", idx, str(dim), metric, qf, str(nq), str(k), str(ef), fi, fd],
                               check=True)
                out[f"ids_ef{ef}"] = np.fromfile(fi, dtype=np.int32).reshape(nq, k)
                out[f"dists_ef{ef}"] = np.fromfile(fd, dtype=np.float32).reshape(nq, k)
            np.savez(os.path.join(HERE, name + ".npz"), **out)
        print("wrote", name)


if __name__ == "__main__":
    main()
