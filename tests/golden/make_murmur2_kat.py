"""Write tests/golden/murmur2_kat.npz: known-answer vectors of the reference's
MurmurHash2 and MurmurHash2_KeyLen4 (unit_tests/bloom_and_hash/sw_murmur2/
murmur2.cpp:8-90), computed by oracle/_ref/ref_murmur2, which links that
source compiled from the reference tree (oracle/Makefile `ref`):

    make -C oracle ref && python tests/golden/make_murmur2_kat.py

Cases: 1,000 seeded (4-byte key, seed) pairs (both functions), and 240
seeded byte strings of length 0..11 (the generic function's tail cases).
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_murmur2")


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run `make -C oracle ref` where the reference tree exists")
    rng = np.random.default_rng(20240)
    keys = rng.integers(0, 2**32, size=1000, dtype=np.uint64).astype(np.uint32)
    keys[:4] = [0, 1, 0xFFFFFFFF, 0x80000000]
    seeds = rng.integers(0, 2**32, size=1000, dtype=np.uint64).astype(np.uint32)
    seeds[:4] = [0, 0x2545F491, 0x2545F492, 0xFFFFFFFF]
    lens = np.repeat(np.arange(12), 20).astype(np.int32)
    blobs = rng.integers(0, 256, size=(lens.size, 12), dtype=np.uint64).astype(np.uint8)
    bseeds = rng.integers(0, 2**32, size=lens.size, dtype=np.uint64).astype(np.uint32)
    lines = []
    for k, s in zip(keys, seeds):
        lines.append(f"4 {int(s)} {np.array([k], dtype='<u4').tobytes().hex()}")
    for n, b, s in zip(lens, blobs, bseeds):
        lines.append(f"{int(n)} {int(s)} {b[:n].tobytes().hex() or '-'}")
    out = subprocess.run([REF], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True).stdout
    vals = np.array([[int(t) for t in ln.split()] for ln in out.splitlines()], dtype=np.uint64).astype(np.uint32)
    assert vals.shape == (len(lines), 2)
    np.savez(os.path.join(HERE, "murmur2_kat.npz"), keys=keys, seeds=seeds, h_generic4=vals[:1000, 0],
             h_keylen4=vals[:1000, 1], lens=lens, blobs=blobs, blob_seeds=bseeds, h_blobs=vals[1000:, 0])
    print("wrote murmur2_kat.npz")


if __name__ == "__main__":
    main()
