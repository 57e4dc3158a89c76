"""Build libgvs.so (HIP kernels for gfx950 + host C++) in-tree.

    python falcon-accelerate-graph-vector-search_amd/build_lib.py [--jobs N]

Host-only sources are compiled with g++ (-ffp-contract=off keeps the HNSW
builder's scalar distance order identical to hnswlib's), device sources with
hipcc for gfx950 only.  Objects go to _build/ (not committed); the shared
library to lib/libgvs.so next to this file, with the C++ host program
lib/gvs_search (cli/gvs_search.cpp) linked against it.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libgvs.so")

HOST_SRCS = ["hnsw_build.cpp", "formats.cpp", "datagen.cpp", "gvs_net.cpp"]
HIP_SRCS = ["gvs_device.hip", "gvs_bruteforce.hip", "gvs_server.hip", "gvs_build.hip",
            "k_f32_l2.hip", "k_f32_ip.hip", "k_bf16_l2.hip", "k_bf16_ip.hip", "k_fp8_l2.hip", "k_fp8_ip.hip",
            "ks_f32_l2.hip", "ks_f32_ip.hip", "k_f32seq_l2.hip", "k_f32seq_ip.hip",
            "ks_bf16_l2.hip", "ks_bf16_ip.hip", "ks_fp8_l2.hip", "ks_fp8_ip.hip",
            "kp_f32_l2.hip", "kp_f32_ip.hip", "kp_bf16_l2.hip", "kp_bf16_ip.hip", "kp_fp8_l2.hip", "kp_fp8_ip.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
HOST_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-ffp-contract=off",
    "-march=x86-64-v3", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(jobs=8, verbose=False, extra_hip_flags=(), out=None, build_dir=None):
    """Compile everything; extra_hip_flags (e.g. -DGVS_MIN_WAVES_PER_EU=5)
    and out/build_dir allow side-by-side variants of the library."""
    bdir = build_dir or BUILD
    lib = out or LIB
    os.makedirs(bdir, exist_ok=True)
    os.makedirs(os.path.dirname(lib), exist_ok=True)
    cmds, objs = [], []
    for s in HOST_SRCS:
        o = os.path.join(bdir, s + ".o")
        cmds.append(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SRCS:
        src = os.path.join(CSRC, s)
        if not os.path.exists(src):
            continue
        o = os.path.join(bdir, s + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, *extra_hip_flags, "-c", src, "-o", o])
        objs.append(o)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for r in ex.map(_run, cmds):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr, file=sys.stderr)
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, lib)
    if out is None:
        # the C++ host program over the C ABI, next to the library
        cli = os.path.join(LIBDIR, "gvs_search")
        _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
              os.path.join(PKG, "cli", "gvs_search.cpp"), "-o", cli + ".tmp", "-L" + LIBDIR, "-lgvs",
              "-Wl,-rpath,$ORIGIN"])
        os.replace(cli + ".tmp", cli)
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    ap.add_argument("--hipflag", action="append", default=[])
    ap.add_argument("--out", default=None)
    ap.add_argument("--build-dir", default=None)
    a = ap.parse_args()
    print(build(a.jobs, a.v, a.hipflag, a.out, a.build_dir))
