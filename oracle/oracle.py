"""ctypes wrapper of the CPU oracle (oracle/liboracle.so): TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, to compute expected answers and the CPU baseline timing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "ref_hnsw")

L2, IP = 0, 1
SEQ, WAVE = 0, 1
F32, BF16, FP8 = 0, 1, 2


class OrcGraph(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("dim", C.c_int), ("dim_pad", C.c_int), ("metric", C.c_int), ("storage", C.c_int),
        ("vec_f32", C.c_void_p), ("vec_bf16", C.c_void_p), ("vec_fp8", C.c_void_p), ("vec_scale", C.c_void_p),
        ("r0", C.c_int), ("adj0", C.c_void_p), ("max_level", C.c_int), ("entry", C.c_int32),
        ("ru", C.c_int), ("upper_off", C.c_void_p), ("levels", C.c_void_p), ("adj_up", C.c_void_p),
    ]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P, I, I64 = C.c_void_p, C.c_int, C.c_int64
    lib.orc_dist_seq.restype = C.c_float
    lib.orc_dist_seq.argtypes = [P, P, I, I]
    lib.orc_dist_wave.restype = C.c_float
    lib.orc_dist_wave.argtypes = [P, P, I, I]
    lib.orc_dist_wave_e.restype = C.c_float
    lib.orc_dist_wave_e.argtypes = [P, P, I, I, I]
    lib.orc_dist_f64.restype = C.c_double
    lib.orc_dist_f64.argtypes = [P, P, I, I]
    lib.orc_dist.restype = C.c_float
    lib.orc_dist.argtypes = [P, P, I64, I]
    lib.orc_search_batch.restype = I
    lib.orc_search_batch.argtypes = [P, P, I64, I, I, I, I, I, P, P, P, P, I]
    lib.orc_bruteforce_f64.restype = None
    lib.orc_bruteforce_f64.argtypes = [P, I64, I, I, P, I64, I, P, P, I]
    lib.orc_murmur2.restype = C.c_uint32
    lib.orc_murmur2.argtypes = [P, I, C.c_uint32]
    lib.orc_murmur2_keylen4.restype = C.c_uint32
    lib.orc_murmur2_keylen4.argtypes = [C.c_uint32, C.c_uint32]
    lib.orc_bloom_test_and_set.restype = I
    lib.orc_bloom_test_and_set.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32]
    lib.orc_bf16_to_f32.restype = C.c_float
    lib.orc_bf16_to_f32.argtypes = [C.c_uint16]
    lib.orc_e4m3_to_f32.restype = C.c_float
    lib.orc_e4m3_to_f32.argtypes = [C.c_uint8]
    lib.orc_f32_to_bf16.restype = C.c_uint16
    lib.orc_f32_to_bf16.argtypes = [C.c_float]
    lib.orc_f32_to_e4m3.restype = C.c_uint8
    lib.orc_f32_to_e4m3.argtypes = [C.c_float]
    lib.orc_search_batch_ex.restype = I
    lib.orc_search_batch_ex.argtypes = [P, P, I64, I, I, I, I, I, P, P, P, P, I, I]
    lib.orc_search_knn_trace.restype = I
    lib.orc_search_knn_trace.argtypes = [P, P, I, I, I, I, I, P, P, P, P, P, P, I, P]
    lib.orc_rerank_f32.restype = I
    lib.orc_rerank_f32.argtypes = [P, I, I, P, P, I, I, I, P, P]
    lib.orc_search_pool.restype = I
    lib.orc_search_pool.argtypes = [P, P, I, I, I, I, I, P, I, P, P, P]
    lib.orc_nsg_search.restype = I
    lib.orc_nsg_search.argtypes = [P, P, I, I, P, I, I, P, P, P]
    return lib


lib = _load()


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def pad_rows(x, dim_pad):
    x = np.asarray(x, dtype=np.float32)
    out = np.zeros((x.shape[0], dim_pad), dtype=np.float32)
    out[:, : x.shape[1]] = x
    return out


class OracleGraph:
    """Holds numpy arrays for the oracle's flat graph (keeps them alive)."""

    def __init__(self, arrays, dim_pad=None, storage=F32, bf16=None, fp8=None, scale=None):
        a = arrays
        dim = a["dim"]
        self.dim_pad = dim_pad or ((dim + 31) // 32 * 32)
        self.vec = pad_rows(a["vectors"], self.dim_pad)
        self.adj0 = np.ascontiguousarray(a["adj0"], dtype=np.int32)
        self.upper_off = np.ascontiguousarray(a["upper_off"], dtype=np.int64)
        self.adj_up = np.ascontiguousarray(a["adj_up"], dtype=np.int32)
        self.bf16, self.fp8, self.scale = bf16, fp8, scale
        g = OrcGraph()
        g.n = a["n"]
        g.dim = dim
        g.dim_pad = self.dim_pad
        g.metric = a["metric"]
        g.storage = storage
        g.vec_f32 = _p(self.vec)
        g.vec_bf16 = _p(bf16)
        g.vec_fp8 = _p(fp8)
        g.vec_scale = _p(scale)
        g.r0 = a["r0"]
        g.adj0 = _p(self.adj0)
        g.max_level = a["max_level"]
        g.entry = a["entry"]
        g.ru = a["ru"]
        g.upper_off = _p(self.upper_off) if a["max_level"] > 0 else None
        g.levels = None
        g.adj_up = _p(self.adj_up) if a["max_level"] > 0 else None
        self.g = g

    def search(self, queries, k, ef, mc=1, mg=1, order=WAVE, entries=None, nthreads=8, with_groups=False):
        """-> ids, dists, stats (hops, evals, upper hops, upper evals[, DST groups])."""
        q = pad_rows(queries, self.dim_pad)
        nq = q.shape[0]
        ids = np.empty((nq, k), dtype=np.int32)
        ds = np.empty((nq, k), dtype=np.float32)
        w = 5 if with_groups else 4
        st = np.empty((nq, w), dtype=np.int64)
        ent = None if entries is None else np.ascontiguousarray(entries, dtype=np.int32)
        lib.orc_search_batch_ex(C.byref(self.g), _p(q), nq, k, ef, mc, mg, order, _p(ent), _p(ids), _p(ds),
                                _p(st), w, nthreads)
        return ids, ds, st

    def trace(self, query, k, ef, mc=1, mg=1, order=WAVE, entry=None, cap=4096):
        """One query: ids, dists, stats (5 columns) and the layer-0 expansion
        order as a list of DST groups, each the list of node ids it expanded."""
        q = pad_rows(np.asarray(query, dtype=np.float32).reshape(1, -1), self.dim_pad)
        ids = np.empty(k, dtype=np.int32)
        ds = np.empty(k, dtype=np.float32)
        st = np.zeros(5, dtype=np.int64)
        tid = np.empty(cap, dtype=np.int32)
        tgr = np.empty(cap, dtype=np.int32)
        nt = C.c_int(0)
        ent = None if entry is None else np.array([entry], dtype=np.int32)
        lib.orc_search_knn_trace(C.byref(self.g), _p(q), k, ef, mc, mg, order, _p(ent), _p(ids), _p(ds), _p(st),
                                 _p(tid), _p(tgr), cap, C.byref(nt))
        groups = []
        for i in range(min(nt.value, cap)):
            if tgr[i] >= len(groups):
                groups.append([])
            groups[tgr[i]].append(int(tid[i]))
        return ids, ds, st, groups


class _Stats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("hops", "evals", "up_hops", "up_evals", "groups")]


def _init_rows(init, nq):
    init = np.ascontiguousarray(init, dtype=np.int32)
    if init.ndim == 1:
        init = np.tile(init, (nq, 1))
    return init


def search_pool(og, queries, k, ef, init, mc=1, mg=1, order=WAVE):
    """searchKnn from an initial pool per query (init: [nq, L] or [L] ids,
    -1 = none) -> ids, dists, stats (hops, evals)."""
    q = pad_rows(queries, og.dim_pad)
    nq = q.shape[0]
    init = _init_rows(init, nq)
    ids = np.empty((nq, k), dtype=np.int32)
    ds = np.empty((nq, k), dtype=np.float32)
    st = np.zeros((nq, 2), dtype=np.int64)
    for i in range(nq):
        s = _Stats()
        lib.orc_search_pool(C.byref(og.g), _p(q[i]), k, ef, mc, mg, order, _p(init[i]), init.shape[1], _p(ids[i]),
                            _p(ds[i]), C.byref(s))
        st[i] = (s.hops, s.evals)
    return ids, ds, st


def nsg_search(og, queries, K, L, init, order=WAVE):
    """nsg.py search_with_base_graph restated on a sorted retset -> ids,
    dists, stats (hops, evals)."""
    q = pad_rows(queries, og.dim_pad)
    nq = q.shape[0]
    init = _init_rows(init, nq)
    ids = np.empty((nq, K), dtype=np.int32)
    ds = np.empty((nq, K), dtype=np.float32)
    st = np.zeros((nq, 2), dtype=np.int64)
    for i in range(nq):
        s = _Stats()
        lib.orc_nsg_search(C.byref(og.g), _p(q[i]), K, L, _p(init[i]), init.shape[1], order, _p(ids[i]), _p(ds[i]),
                           C.byref(s))
        st[i] = (s.hops, s.evals)
    return ids, ds, st


def bruteforce_f64(base, queries, k, metric=L2, nthreads=8):
    base = np.ascontiguousarray(base, dtype=np.float32)
    q = np.ascontiguousarray(queries, dtype=np.float32)
    ids = np.empty((q.shape[0], k), dtype=np.int32)
    ds = np.empty((q.shape[0], k), dtype=np.float64)
    lib.orc_bruteforce_f64(_p(base), base.shape[0], base.shape[1], metric, _p(q), q.shape[0], k, _p(ids),
                           _p(ds), nthreads)
    return ids, ds


def recall_at(found, truth, k):
    """Recall@k as the reference computes it: |top-k found ∩ top-k truth| / k
    (FPGA_multi_DDR/FPGA_inter_query_v1.3_longer_FIFO_alt_PR/src/host.cpp:1051-1063)."""
    found = np.asarray(found)[:, :k]
    truth = np.asarray(truth)[:, :k]
    hit = sum(len(set(f.tolist()) & set(t.tolist())) for f, t in zip(found, truth))
    return hit / float(truth.shape[0] * k)


def rerank_f32(og, queries, cand, k, order=WAVE):
    """Exact fp32 rerank of candidate ids per query (oracle.c orc_rerank_f32)."""
    q = pad_rows(queries, og.dim_pad)
    cand = np.ascontiguousarray(cand, dtype=np.int32)
    ids = np.empty((q.shape[0], k), dtype=np.int32)
    ds = np.empty((q.shape[0], k), dtype=np.float32)
    for i in range(q.shape[0]):
        lib.orc_rerank_f32(_p(og.vec), og.dim_pad, og.g.metric, _p(q[i]), _p(cand[i]), cand.shape[1], k, order,
                           _p(ids[i]), _p(ds[i]))
    return ids, ds
