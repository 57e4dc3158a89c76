"""ctypes binding of libgvs.so (the C ABI in include/gvs.h).

Mirrors the Python surface of hnswlib (``Index(space, dim)``,
``init_index``/``add_items`` (build), ``set_ef``, ``knn_query``,
``load_index``/``save_index``) that the reference's scripts drive
(vector_search_baselines/scripts_hnsw/construct_and_search_hnsw.py:126-148,
vector_search_baselines/scripts_hnsw/hnsw.py:379-517), with the search running
on the GPU through the HIP library.  There is no CPU search fallback: if the
library is missing, importing the module raises.
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("GVS_LIBRARY") or os.path.join(_PKG, "lib", "libgvs.so")

GVS_OK = 0
FLAG_ASYNC = 1
FLAG_HBM_VISITED = 2
FLAG_LDS_VISITED = 4
FLAG_LDS_WIDE = 32
FLAG_REGION_ORDER = 8
FLAG_HNSWLIB_ORDER = 16
ERRORS = {-1: "ARG", -2: "HIP", -3: "NOMEM", -4: "UNSUPPORTED", -5: "OVERFLOW", -6: "IO", -7: "STATE"}
L2, IP = 0, 1
F32, BF16, FP8 = 0, 1, 2
STATS_PER_QUERY = 4


class GvsError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"gvs error {rc} ({ERRORS.get(rc, '?')}): {msg}")
        self.rc = rc


class Graph(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("dim", C.c_int32), ("metric", C.c_int32),
        ("vectors", C.c_void_p), ("ld_vectors", C.c_int64),
        ("r0", C.c_int32), ("adj0", C.c_void_p),
        ("max_level", C.c_int32), ("entry", C.c_int32),
        ("ru", C.c_int32), ("up_rows", C.c_int64),
        ("upper_off", C.c_void_p), ("adj_up", C.c_void_p), ("labels", C.c_void_p),
    ]


class Params(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "k", "ef", "mc", "mg", "waves_per_query", "storage", "rerank", "persistent",
        "visited_log2", "bloom_log2", "grid_waves", "flags")]


class ServerConfig(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "max_inflight", "max_k", "max_ef", "queue_bytes", "grid_waves", "mc", "storage", "waves_per_request",
        "rerank")]


class BuildConfig(C.Structure):
    _fields_ = [("exact_order", C.c_int32), ("max_batch", C.c_int32), ("host_threads", C.c_int32),
                ("reserved", C.c_int32), ("batch_ratio", C.c_double)]


class NetConfig(C.Structure):
    _fields_ = [("c2f_port", C.c_int32), ("f2c_host", C.c_char_p), ("f2c_port", C.c_int32), ("dim", C.c_int32),
                ("ef", C.c_int32), ("timeout_ms", C.c_int32)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    P, I32, I64, U32, U64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64
    sig = {
        "gvs_last_error": (C.c_char_p, []),
        "gvs_version": (C.c_char_p, []),
        "gvs_hnsw_build": (C.c_int, [P, I64, I32, I32, I32, I32, U64, I32, I32, P]),
        "gvs_nsg_from_hnsw": (C.c_int, [P, I32, P]),
        "gvs_hnswlib_load": (C.c_int, [C.c_char_p, I32, I32, P]),
        "gvs_hnswlib_save": (C.c_int, [P, C.c_char_p]),
        "gvs_nsg_load": (C.c_int, [C.c_char_p, P, I64, I32, I32, P]),
        "gvs_fpga_load": (C.c_int, [C.c_char_p, I32, I32, I32, I32, P]),
        "gvs_fpga_save": (C.c_int, [P, C.c_char_p, I32]),
        "gvs_host_graph_view": (C.c_int, [P, P]),
        "gvs_host_graph_free": (None, [P]),
        "gvs_gen_mixture": (C.c_int, [U64, U64, I64, I32, I32, C.c_float, C.c_float, I32, P, I32]),
        "gvs_index_create": (C.c_int, [P, U32, P]),
        "gvs_index_destroy": (C.c_int, [P]),
        "gvs_index_fork": (C.c_int, [P, P]),
        "gvs_index_device_bytes": (C.c_int, [P, P, P]),
        "gvs_search_device": (C.c_int, [P, P, I64, I64, P, P, P, P, P, P]),
        "gvs_search_host": (C.c_int, [P, P, I64, I64, P, P, P, P, P]),
        "gvs_search_device_noalloc": (C.c_int, [P, P, I64, I64, P, P, P, P, P, P]),
        "gvs_search_overflow_count": (C.c_int, [P, P]),
        "gvs_search_counters": (C.c_int, [P, P]),
        "gvs_autotune": (C.c_int, [P, P, I64, I64, P, P, P, P, P, P]),
        "gvs_search_graph_create": (C.c_int, [P, P, I64, I64, P, P, P, P, P, P]),
        "gvs_search_graph_launch": (C.c_int, [P, P]),
        "gvs_search_graph_destroy": (C.c_int, [P]),
        "gvs_device_copy": (C.c_int, [P, P, I64, P]),
        "gvs_stats_per_query": (C.c_int, []),
        "gvs_search_plan": (C.c_int, [P, I64, P, P]),
        "gvs_hash_murmur2_keylen4_device": (C.c_int, [P, P, P, I64, P]),
        "gvs_bruteforce_f64_device": (C.c_int, [P, P, I64, I64, I32, P, P, P]),
        "gvs_server_open": (C.c_int, [P, P, P]),
        "gvs_server_submit": (I64, [P, P, I32, I32]),
        "gvs_server_submit_batch": (I64, [P, P, I64, I64, P, P, P]),
        "gvs_server_poll_status": (C.c_int, [P, I32, P, P, P, P, P, P]),
        "gvs_server_counters": (C.c_int, [P, P, I32]),
        "gvs_server_loadgen_breakdown": (C.c_int, [P, P]),
        "gvs_server_poll": (C.c_int, [P, I32, P, P, P, P]),
        "gvs_server_cancel": (C.c_int, [P, I64]),
        "gvs_server_poll_stats": (C.c_int, [P, I32, P, P, P, P, P]),
        "gvs_server_loadgen": (C.c_int, [P, P, I64, I64, C.c_double, U64, P, P, P, P, P, P]),
        "gvs_server_loadgen_mt": (C.c_int, [P, P, I64, I64, C.c_double, U64, P, P, P, P, P, P, I32, I32]),
        "gvs_server_stats": (C.c_int, [P, P]),
        "gvs_server_close": (C.c_int, [P]),
        "gvs_net_serve": (C.c_int, [P, P, P, P]),
        "gvs_hnsw_build_device": (C.c_int, [P, I64, I32, I32, I32, I32, U64, P, P, P]),
        "gvs_hnsw_build_batched": (C.c_int, [P, I64, I32, I32, I32, I32, U64, I32, I32, C.c_double, I32, P]),
        "gvs_search_pool_device": (C.c_int, [P, P, I64, I64, P, I32, I64, P, P, P, P, P]),
        "gvs_search_pool_host": (C.c_int, [P, P, I64, I64, P, I32, I64, P, P, P, P]),
        "gvs_net_frame_bytes": (C.c_int, [I32, I32, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()
# stats row width of the loaded library (4 counters; 16 when compiled with phase profiling)
STATS_ROW = lib.gvs_stats_per_query()


def exported_symbols():
    """Names of the C ABI functions this binding knows, present in the library."""
    names = ["gvs_last_error", "gvs_version", "gvs_hnsw_build", "gvs_nsg_from_hnsw", "gvs_hnswlib_load",
             "gvs_hnswlib_save", "gvs_nsg_load", "gvs_fpga_load", "gvs_fpga_save", "gvs_host_graph_view", "gvs_host_graph_free",
             "gvs_gen_mixture", "gvs_index_create", "gvs_index_destroy", "gvs_index_fork", "gvs_index_device_bytes",
             "gvs_search_device", "gvs_search_host", "gvs_search_device_noalloc", "gvs_search_overflow_count", "gvs_search_counters", "gvs_autotune", "gvs_search_graph_create",
             "gvs_search_graph_launch", "gvs_search_graph_destroy", "gvs_device_copy", "gvs_stats_per_query", "gvs_search_plan", "gvs_hash_murmur2_keylen4_device", "gvs_bruteforce_f64_device", "gvs_server_open",
             "gvs_server_submit", "gvs_server_submit_batch", "gvs_server_poll", "gvs_server_poll_stats",
             "gvs_server_poll_status", "gvs_server_counters", "gvs_server_loadgen_breakdown", "gvs_server_cancel",
             "gvs_server_stats", "gvs_server_close", "gvs_server_loadgen", "gvs_server_loadgen_mt", "gvs_net_serve", "gvs_net_frame_bytes", "gvs_search_pool_device", "gvs_search_pool_host", "gvs_hnsw_build_device", "gvs_hnsw_build_batched"]
    return [n for n in names if hasattr(lib, n)]


def _check(rc):
    if rc != GVS_OK:
        raise GvsError(rc, lib.gvs_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_copy(dst_ptr, src_ptr, nbytes, stream=None):
    """Device-to-device copy through the library's streaming kernel."""
    _check(lib.gvs_device_copy(C.c_void_p(dst_ptr), C.c_void_p(src_ptr), nbytes, C.c_void_p(stream) if stream else None))


def murmur2_keylen4_device(keys_ptr, seeds_ptr, out_ptr, n, stream=None):
    """The device Bloom hash of n device words (MurmurHash2_KeyLen4)."""
    _check(lib.gvs_hash_murmur2_keylen4_device(C.c_void_p(keys_ptr), C.c_void_p(seeds_ptr), C.c_void_p(out_ptr), n,
                                               C.c_void_p(stream) if stream else None))


def gen_mixture(n, dim, ncenters, seed_centers, seed_points, center_std=4.0, noise_std=1.0,
                normalize=False, nthreads=8):
    """Seeded Gaussian mixture (float32, n x dim)."""
    out = np.empty((n, dim), dtype=np.float32)
    _check(lib.gvs_gen_mixture(seed_centers, seed_points, n, dim, ncenters, center_std, noise_std,
                               int(bool(normalize)), _ptr(out), nthreads))
    return out


class HostGraph:
    """Owned host graph (HNSW levels + base layer, or flat)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if isinstance(handle, int) else handle
        self._view = Graph()
        _check(lib.gvs_host_graph_view(self._h, C.byref(self._view)))

    @classmethod
    def build_hnsw(cls, x, M=16, ef_construction=200, seed=100, metric=L2, nthreads=8, exact_order=False):
        x = np.ascontiguousarray(x, dtype=np.float32)
        h = C.c_void_p()
        _check(lib.gvs_hnsw_build(_ptr(x), x.shape[0], x.shape[1], metric, M, ef_construction, seed,
                                  nthreads, int(bool(exact_order)), C.byref(h)))
        return cls(h)

    @classmethod
    def build_hnsw_gpu(cls, x, M=16, ef_construction=200, seed=100, metric=L2, exact_order=False,
                       batch_ratio=0.0, max_batch=0, host_threads=0, with_stats=False):
        """HNSW built on the GPU (gvs_hnsw_build_device); batch_ratio < 0 =
        batches of one node (hnswlib's sequential insertion)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        h = C.c_void_p()
        cfg = BuildConfig(int(bool(exact_order)), max_batch, host_threads, 0, batch_ratio)
        st = (C.c_double * 8)()
        _check(lib.gvs_hnsw_build_device(_ptr(x), x.shape[0], x.shape[1], metric, M, ef_construction, seed,
                                         C.byref(cfg), C.byref(h), st))
        g = cls(h)
        if with_stats:
            return g, dict(zip(("seconds", "upper_s", "search_s", "link_s", "batches"), list(st)[:5]))
        return g

    @classmethod
    def build_hnsw_batched(cls, x, M=16, ef_construction=200, seed=100, metric=L2, nthreads=8, exact_order=False,
                           batch_ratio=0.0, max_batch=0):
        """The GPU builder's batched insertion on host threads
        (gvs_hnsw_build_batched)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        h = C.c_void_p()
        _check(lib.gvs_hnsw_build_batched(_ptr(x), x.shape[0], x.shape[1], metric, M, ef_construction, seed,
                                          nthreads, int(bool(exact_order)), batch_ratio, max_batch, C.byref(h)))
        return cls(h)

    @classmethod
    def load_hnswlib(cls, path, dim, metric=L2):
        h = C.c_void_p()
        _check(lib.gvs_hnswlib_load(path.encode(), dim, metric, C.byref(h)))
        return cls(h)

    @classmethod
    def load_nsg(cls, path, x, metric=L2):
        x = np.ascontiguousarray(x, dtype=np.float32)
        h = C.c_void_p()
        _check(lib.gvs_nsg_load(path.encode(), _ptr(x), x.shape[0], x.shape[1], metric, C.byref(h)))
        return cls(h)

    @classmethod
    def load_fpga(cls, directory, kind, dim, nchannels=1, metric=L2):
        """The reference's FPGA-layout directory (kind 0 = HNSW, 1 = NSG)."""
        h = C.c_void_p()
        _check(lib.gvs_fpga_load(str(directory).encode(), kind, dim, nchannels, metric, C.byref(h)))
        return cls(h)

    def save_fpga(self, directory, nchannels=1):
        _check(lib.gvs_fpga_save(self._h, str(directory).encode(), nchannels))

    def to_nsg(self, R):
        h = C.c_void_p()
        _check(lib.gvs_nsg_from_hnsw(self._h, R, C.byref(h)))
        return HostGraph(h)

    def save_hnswlib(self, path):
        _check(lib.gvs_hnswlib_save(self._h, path.encode()))

    @property
    def view(self):
        return self._view

    def arrays(self):
        """numpy copies of the flat arrays (for the oracle and tests)."""
        v = self._view

        def arr(ptr, count, dt):
            if not ptr or count == 0:
                return np.zeros(0, dtype=dt)
            buf = (C.c_char * (count * np.dtype(dt).itemsize)).from_address(ptr)
            return np.frombuffer(buf, dtype=dt).copy()

        out = {
            "n": v.n, "dim": v.dim, "metric": v.metric, "r0": v.r0, "max_level": v.max_level,
            "entry": v.entry, "ru": v.ru,
            "vectors": arr(v.vectors, v.n * v.dim, np.float32).reshape(v.n, v.dim),
            "adj0": arr(v.adj0, v.n * v.r0, np.int32).reshape(v.n, v.r0),
            "upper_off": arr(v.upper_off, v.n if v.max_level > 0 else 0, np.int64),
            "adj_up": arr(v.adj_up, v.up_rows * v.ru, np.int32).reshape(-1, max(v.ru, 1)),
            "labels": arr(v.labels, v.n, np.int64),
        }
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and lib is not None:
            lib.gvs_host_graph_free(h)
            self._h = None


def make_params(k=10, ef=64, mc=1, mg=1, waves_per_query=1, storage=F32, rerank=False, persistent=False,
                visited_log2=0, bloom_log2=0, grid_waves=0, asynchronous=False, visited=None, region_order=False,
                hnswlib_order=False):
    """visited: None (auto), "hbm" (LDS Bloom + HBM hash), "lds" (exact LDS
    hash; 2-byte slots when max(ef, k) > 128 and the graph's ids fit them;
    visited_log2 in [10, 14] then sets its size) or "lds4" (the same with
    4-byte slots at every ef).
    region_order: run the batch in region order (persistent mode; GVS_FLAG_REGION_ORDER).
    hnswlib_order: sum distances in hnswlib's scalar order (GVS_FLAG_HNSWLIB_ORDER)."""
    flags = FLAG_ASYNC if asynchronous else 0
    if region_order:
        flags |= FLAG_REGION_ORDER
    if hnswlib_order:
        flags |= FLAG_HNSWLIB_ORDER
    if visited == "hbm":
        flags |= FLAG_HBM_VISITED
    elif visited == "lds":
        flags |= FLAG_LDS_VISITED
    elif visited == "lds4":
        flags |= FLAG_LDS_VISITED | FLAG_LDS_WIDE
    return Params(k, ef, mc, mg, waves_per_query, storage, int(bool(rerank)), int(bool(persistent)),
                  visited_log2, bloom_log2, grid_waves, flags)


class DeviceIndex:
    """Graph + vectors resident on the current HIP device."""

    def __init__(self, graph, storages=(F32,)):
        mask = 0
        for s in storages:
            mask |= 1 << s
        self._graph = graph  # keep host arrays alive for the view
        self._h = C.c_void_p()
        _check(lib.gvs_index_create(C.byref(graph.view), mask, C.byref(self._h)))
        self.dim = graph.view.dim
        self.n = graph.view.n

    @property
    def handle(self):
        return self._h

    def fork(self):
        """A second handle on the same device arrays with scratch of its own
        (gvs_index_fork): searches through different handles may run at once
        on different streams."""
        f = DeviceIndex.__new__(DeviceIndex)
        f._graph = self._graph
        f._root = getattr(self, "_root", None) or self  # destroyed after its forks
        f._h = C.c_void_p()
        _check(lib.gvs_index_fork(self._h, C.byref(f._h)))
        f.dim, f.n = self.dim, self.n
        return f

    def search(self, queries, k=10, ef=64, entries=None, with_stats=False, **kw):
        """searchKnn over a batch of host queries -> (labels, dists[, stats])."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 1:
            q = q[None, :]
        nq = q.shape[0]
        lab = np.empty((nq, k), dtype=np.int64)
        dist = np.empty((nq, k), dtype=np.float32)
        st = np.empty((nq, STATS_ROW), dtype=np.int32) if with_stats else None
        ent = None if entries is None else np.ascontiguousarray(entries, dtype=np.int32)
        p = make_params(k=k, ef=ef, **kw)
        _check(lib.gvs_search_host(self._h, _ptr(q), nq, q.shape[1], _ptr(ent), C.byref(p), _ptr(lab),
                                   _ptr(dist), _ptr(st)))
        return (lab, dist, st[:, :STATS_PER_QUERY]) if with_stats else (lab, dist)

    def search_pool(self, queries, k=10, ef=64, init=None, with_stats=False, **kw):
        """NSG search from an initial pool per query (init: [nq, L] or [L]
        ids, -1 = none; None = the entry's first neighbours)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 1:
            q = q[None, :]
        nq = q.shape[0]
        ini, ln, ld = None, 0, 0
        if init is not None:
            ini = np.ascontiguousarray(init, dtype=np.int32)
            if ini.ndim == 1:
                ini = np.ascontiguousarray(np.tile(ini, (nq, 1)))
            ln = ld = ini.shape[1]
        lab = np.empty((nq, k), dtype=np.int64)
        dist = np.empty((nq, k), dtype=np.float32)
        st = np.empty((nq, STATS_ROW), dtype=np.int32) if with_stats else None
        p = make_params(k=k, ef=ef, **kw)
        _check(lib.gvs_search_pool_host(self._h, _ptr(q), nq, q.shape[1], _ptr(ini), ln, ld, C.byref(p), _ptr(lab),
                                        _ptr(dist), _ptr(st)))
        return (lab, dist, st[:, :STATS_PER_QUERY]) if with_stats else (lab, dist)

    def search_pool_device(self, q_ptr, nq, ldq, labels_ptr, dists_ptr, k=10, ef=64, init_ptr=None, init_len=0,
                           init_ld=0, stats_ptr=None, stream=None, **kw):
        """gvs_search_pool_device with device buffers (init_ptr None = the
        entry's first neighbours)."""
        p = make_params(k=k, ef=ef, **kw)
        _check(lib.gvs_search_pool_device(self._h, C.c_void_p(q_ptr), nq, ldq,
                                          C.c_void_p(init_ptr) if init_ptr else None, init_len, init_ld,
                                          C.byref(p), C.c_void_p(labels_ptr), C.c_void_p(dists_ptr),
                                          C.c_void_p(stats_ptr) if stats_ptr else None,
                                          C.c_void_p(stream) if stream else None))

    def search_device(self, q_ptr, nq, ldq, labels_ptr, dists_ptr, k=10, ef=64, entries_ptr=None,
                      stats_ptr=None, stream=None, **kw):
        """Search with device pointers (e.g. torch tensors' data_ptr())."""
        p = make_params(k=k, ef=ef, **kw)
        _check(lib.gvs_search_device(self._h, C.c_void_p(q_ptr), nq, ldq,
                                     C.c_void_p(entries_ptr) if entries_ptr else None, C.byref(p),
                                     C.c_void_p(labels_ptr), C.c_void_p(dists_ptr),
                                     C.c_void_p(stats_ptr) if stats_ptr else None,
                                     C.c_void_p(stream) if stream else None))

    def search_device_noalloc(self, q_ptr, nq, ldq, labels_ptr, dists_ptr, k=10, ef=64, entries_ptr=None,
                              stats_ptr=None, stream=None, **kw):
        """gvs_search_device_noalloc: the same search with no allocation at call time."""
        p = make_params(k=k, ef=ef, **kw)
        _check(lib.gvs_search_device_noalloc(self._h, C.c_void_p(q_ptr), nq, ldq,
                                             C.c_void_p(entries_ptr) if entries_ptr else None, C.byref(p),
                                             C.c_void_p(labels_ptr), C.c_void_p(dists_ptr),
                                             C.c_void_p(stats_ptr) if stats_ptr else None,
                                             C.c_void_p(stream) if stream else None))

    def plan(self, nq, k=10, ef=64, **kw):
        """The launch plan of a search of nq queries (gvs_search_plan)."""
        p = make_params(k=k, ef=ef, **kw)
        out = (C.c_int32 * 8)()
        _check(lib.gvs_search_plan(self._h, nq, C.byref(p), out))
        return dict(zip(("lds_hash_log2", "lds_hash_2byte", "hbm_table_log2", "bloom_log2", "lds_bytes",
                         "waves_per_block", "resident_waves", "list_capacity"), list(out)))

    def device_bytes(self):
        """Bytes resident on the device: {"graph": ..., "vectors": ...}."""
        gb, vb = C.c_int64(0), C.c_int64(0)
        _check(lib.gvs_index_device_bytes(self._h, C.byref(gb), C.byref(vb)))
        return {"graph": gb.value, "vectors": vb.value}

    def autotune(self, q_ptr, nq, ldq, labels_ptr, dists_ptr, k=10, ef=64, stream=None, **kw):
        """Fastest launch configuration for this batch -> (config dict, ms per launch)."""
        p = make_params(k=k, ef=ef, **kw)
        best = Params()
        ms = C.c_float(0.0)
        _check(lib.gvs_autotune(self._h, C.c_void_p(q_ptr), nq, ldq, C.byref(p), C.c_void_p(labels_ptr),
                                C.c_void_p(dists_ptr), C.c_void_p(stream) if stream else None, C.byref(best),
                                C.byref(ms)))
        cfg = dict(ef=best.ef, mc=best.mc, mg=best.mg, persistent=bool(best.persistent),
                   waves_per_query=best.waves_per_query, grid_waves=best.grid_waves,
                   region_order=bool(best.flags & FLAG_REGION_ORDER))
        if best.flags & FLAG_HBM_VISITED:
            cfg["visited"] = "hbm"
        return cfg, ms.value

    def capture(self, q_ptr, nq, ldq, labels_ptr, dists_ptr, k=10, ef=64, entries_ptr=None, stats_ptr=None, **kw):
        """Record the search of these device buffers into a hipGraph (SearchGraph)."""
        p = make_params(k=k, ef=ef, **kw)
        h = C.c_void_p()
        _check(lib.gvs_search_graph_create(self._h, C.c_void_p(q_ptr), nq, ldq,
                                           C.c_void_p(entries_ptr) if entries_ptr else None, C.byref(p),
                                           C.c_void_p(labels_ptr), C.c_void_p(dists_ptr),
                                           C.c_void_p(stats_ptr) if stats_ptr else None, C.byref(h)))
        return SearchGraph(h, self)

    def overflow_count(self):
        c = C.c_int64(0)
        _check(lib.gvs_search_overflow_count(self._h, C.byref(c)))
        return c.value

    def counters(self):
        """Device counters of the last search: claimed, overflowed, LDS-hash spills."""
        c = (C.c_int64 * 3)()
        _check(lib.gvs_search_counters(self._h, c))
        return {"claimed": c[0], "overflowed": c[1], "spilled": c[2]}

    def bruteforce_f64_device(self, q_ptr, nq, ldq, k, ids_ptr, dists_ptr, stream=None):
        _check(lib.gvs_bruteforce_f64_device(self._h, C.c_void_p(q_ptr), nq, ldq, k, C.c_void_p(ids_ptr),
                                             C.c_void_p(dists_ptr), C.c_void_p(stream) if stream else None))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and lib is not None:
            lib.gvs_index_destroy(h)
            self._h = None


class SearchGraph:
    """A captured search (C ABI gvs_search_graph_*): launch() replays it."""

    def __init__(self, h, dev):
        self._h = h
        self.dev = dev  # keeps the index alive

    def launch(self, stream=None):
        _check(lib.gvs_search_graph_launch(self._h, C.c_void_p(stream) if stream else None))

    def close(self):
        if self._h is not None and self._h.value:
            _check(lib.gvs_search_graph_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching serving loop over a DeviceIndex (C ABI gvs_server_*)."""

    def __init__(self, dev, max_inflight=1024, max_k=100, max_ef=512, queue_bytes=0, grid_waves=0, mc=1,
                 waves_per_request=1, storage=F32, rerank=False):
        self.dev = dev
        self.max_k = max_k
        cfg = ServerConfig(max_inflight, max_k, max_ef, queue_bytes, grid_waves, mc, storage, waves_per_request,
                           int(bool(rerank)))
        self._h = C.c_void_p()
        _check(lib.gvs_server_open(dev.handle, C.byref(cfg), C.byref(self._h)))

    def submit(self, query, k, ef):
        q = np.ascontiguousarray(query, dtype=np.float32)
        r = lib.gvs_server_submit(self._h, _ptr(q), k, ef)
        if r < 0:
            raise GvsError(int(r), lib.gvs_last_error().decode())
        return r

    def try_submit(self, query, k, ef):
        """Like submit, but returns None when all device slots are busy or the
        queue memory is taken."""
        q = np.ascontiguousarray(query, dtype=np.float32)
        r = lib.gvs_server_submit(self._h, _ptr(q), k, ef)
        if r == -7:
            return None
        if r < 0:
            raise GvsError(int(r), lib.gvs_last_error().decode())
        return r

    def submit_batch(self, queries, ks, efs):
        """Admit as many requests as slots allow, in order, published together
        -> request ids of the admitted prefix."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        ks = np.ascontiguousarray(ks, dtype=np.int32)
        efs = np.ascontiguousarray(efs, dtype=np.int32)
        ids = np.empty(q.shape[0], dtype=np.int64)
        r = lib.gvs_server_submit_batch(self._h, _ptr(q), q.shape[0], q.shape[1], _ptr(ks), _ptr(efs), _ptr(ids))
        if r < 0:
            raise GvsError(int(r), lib.gvs_last_error().decode())
        return ids[:r]

    def poll(self, max_done=256, with_stats=False, with_status=False):
        rid = np.empty(max_done, dtype=np.int64)
        lab = np.empty((max_done, self.max_k), dtype=np.int64)
        dist = np.empty((max_done, self.max_k), dtype=np.float32)
        lat = np.empty(max_done, dtype=np.float64)
        st = np.empty((max_done, STATS_ROW), dtype=np.int32) if with_stats else None
        status = np.empty(max_done, dtype=np.int32) if with_status else None
        n = lib.gvs_server_poll_status(self._h, max_done, _ptr(rid), _ptr(lab), _ptr(dist), _ptr(lat), _ptr(st),
                                       _ptr(status))
        if n < 0:
            raise GvsError(n, lib.gvs_last_error().decode())
        out = (rid[:n], lab[:n], dist[:n], lat[:n])
        if with_stats:
            out = out + (st[:n, :STATS_PER_QUERY],)
        if with_status:
            out = out + (status[:n],)
        return out

    def cancel(self, req):
        return lib.gvs_server_cancel(self._h, req) == GVS_OK

    def stats(self):
        c = np.zeros(7, dtype=np.int64)
        n = lib.gvs_server_counters(self._h, _ptr(c), 7)
        if n < 0:
            raise GvsError(n, lib.gvs_last_error().decode())
        return dict(zip(("submitted", "admitted", "completed", "cancelled", "rejected", "device_loops",
                         "overflowed"), c.tolist()))

    def loadgen(self, queries, rate_qps, ks, efs, seed=1, submit_threads=1, poll_threads=1):
        """Open-loop Poisson load (gvs_server_loadgen_mt) -> (labels, dists,
        latency_us per query, wall seconds)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        ks = np.ascontiguousarray(ks, dtype=np.int32)
        efs = np.ascontiguousarray(efs, dtype=np.int32)
        lab = np.empty((nq, self.max_k), dtype=np.int64)
        dist = np.empty((nq, self.max_k), dtype=np.float32)
        lat = np.empty(nq, dtype=np.float64)
        wall = C.c_double(0.0)
        _check(lib.gvs_server_loadgen_mt(self._h, _ptr(q), nq, q.shape[1], float(rate_qps), seed, _ptr(ks),
                                         _ptr(efs), _ptr(lab), _ptr(dist), _ptr(lat), C.byref(wall),
                                         int(submit_threads), int(poll_threads)))
        return lab, dist, lat, wall.value

    def loadgen_breakdown(self):
        """Host-side time split of the last loadgen (gvs_server_loadgen_breakdown)."""
        out = np.zeros(6, dtype=np.float64)
        _check(lib.gvs_server_loadgen_breakdown(self._h, _ptr(out)))
        return dict(zip(("submit_ns_per_request", "poll_ns_per_request", "submit_busy", "poll_busy",
                         "requests_per_batch", "completions_per_poll"), [round(float(v), 3) for v in out]))

    def close(self):
        if self._h is not None and self._h.value:
            _check(lib.gvs_server_close(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Index:
    """hnswlib-style front end: Index(space, dim) -> init_index / add_items /
    set_ef / knn_query, with the search on the GPU."""

    def __init__(self, space, dim):
        self.space = space
        self.dim = dim
        self.metric = L2 if space == "l2" else IP
        self.ef = 10
        self._graph = None
        self._dev = None
        self._params = {}

    def init_index(self, max_elements, M=16, ef_construction=200, random_seed=100):
        self._params = dict(M=M, ef_construction=ef_construction, seed=random_seed)

    def add_items(self, data, num_threads=8, exact_order=False):
        self._graph = HostGraph.build_hnsw(data, metric=self.metric, nthreads=num_threads,
                                           exact_order=exact_order, **self._params)
        self._dev = None

    def load_index(self, path):
        self._graph = HostGraph.load_hnswlib(path, self.dim, self.metric)
        self._dev = None

    def save_index(self, path):
        self._graph.save_hnswlib(path)

    def set_ef(self, ef):
        self.ef = ef

    def knn_query(self, data, k=1, **kw):
        if self._dev is None:
            self._dev = DeviceIndex(self._graph)
        return self._dev.search(data, k=k, ef=self.ef, **kw)
