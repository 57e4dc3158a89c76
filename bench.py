#!/usr/bin/env python3
"""Benchmark: QPS at recall@10 >= 0.95 on a 1M x 128 HNSW index, one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (synthetic; the reference's SIFT/Deep datasets are not available):
1,000,000 base vectors, D=128, from a seeded 256-centre Gaussian mixture
(centres ~ N(0, 16 I), noise N(0, I)); HNSW M=32 (layer-0 degree 64),
efConstruction=200; 10,000 queries from the same mixture; k=10, L2.
Ground truth: exact FP64 brute force on the GPU (MFMA kernel).

A step is one search of the whole 10k-query batch with queries, graph and
vectors resident in HBM.  ef is the smallest of a fixed sweep that reaches
recall@10 >= 0.95 for the plain one-wavefront-per-query search; the tuned
path (persistent work-queue kernel, DST candidates per step) is chosen by the
autotuner among configurations that keep recall@10 >= 0.95.  rank 0 prints
one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "falcon-accelerate-graph-vector-search_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "QPS @ recall@10≥0.95, 1M×128d HNSW; p50/p99 latency under load vs plain GPU"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def recall_at(found, truth, k):
    f = np.asarray(found)[:, :k]
    t = np.asarray(truth)[:, :k]
    return float((f[:, :, None] == t[:, None, :]).any(-1).sum()) / (t.shape[0] * k)


def read_traffic(path, cfg):
    """HBM bytes per search launch from a committed rocprofv3 PMC summary:
    the line 'traffic ef=E mc=C mg=G persistent=P bytes=B' for this config."""
    if not os.path.exists(path):
        return None
    key = f"ef={cfg['ef']} mc={cfg['mc']} mg={cfg.get('mg', 1)} persistent={int(bool(cfg['persistent']))}"
    if cfg.get("grid_waves", 0):
        key += f" grid_waves={cfg['grid_waves']}"
    if cfg.get("region_order"):
        key += " region_order=1"
    if cfg.get("waves_per_query", 1) > 1:
        key += f" waves_per_query={cfg['waves_per_query']}"
    for line in open(path):
        line = line.strip()
        if line.startswith("traffic ") and " bytes=" in line and line.split(" bytes=")[0] == "traffic " + key:
            return int(line.split(" bytes=")[1].split()[0])
    return None


def host_cpu():
    """Host CPU model, logical CPU count and the CPUs this process may use."""
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    allowed = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    return model, os.cpu_count() or 1, allowed


def cpu_baseline(a, g, q, gt, ef, index_path):
    """The reference's own hnswlib searchKnn (oracle/_ref/ref_hnsw_simd:
    compiled from the reference's headers with its SIMD distance kernels,
    -O3 -march=x86-64-v3, OpenMP across queries), pinned with taskset to this
    process's host CPUs (at most 16, the GPU's share of the host), on the
    same index and queries as the GPU run, for about cpu_seconds; plus batch=1
    latency on one thread.  Where that binary is not built (no reference
    tree where the library was built), the hnswlib-order oracle stands in
    (kind "port")."""
    import subprocess
    import tempfile
    model, nproc, allowed = host_cpu()
    cpus = allowed[:16]
    threads = len(cpus)
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_hnsw_simd")
    if os.path.exists(ref) and os.path.exists(index_path):
        with tempfile.TemporaryDirectory() as td:
            qf, idf = os.path.join(td, "q.f32"), os.path.join(td, "ids.i32")
            np.ascontiguousarray(q, dtype=np.float32).tofile(qf)
            cmd = ["taskset", "-c", ",".join(str(c) for c in cpus), ref, "bench", index_path, str(a.dim), "l2", qf,
                   str(len(q)), str(a.k), str(ef), str(threads), str(a.cpu_seconds), idf]
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
            if r.returncode == 0:
                res = json.loads(r.stdout.strip().splitlines()[-1])
                ids = np.fromfile(idf, dtype=np.int32).reshape(len(q), a.k)
                cpu = {"value": res["qps"], "unit": "QPS", "cores": threads, "kind": "reference",
                       "sample": f"{res['passes']} pass(es) over the {len(q)} queries ({res['seconds']} s), ef={ef}, "
                                 f"k={a.k}: the reference's hnswlib searchKnn (oracle/_ref/ref_hnsw_simd, SIMD "
                                 f"-O3 -march=x86-64-v3, OpenMP over queries) on the GPU run's index, taskset to "
                                 f"{threads} host CPUs",
                       "recall_at_10": round(recall_at(ids, gt, a.k), 4), "cpu_model": model, "nproc": nproc,
                       "batch1_latency_us": {"p50": res["batch1_p50_us"], "p99": res["batch1_p99_us"],
                                             "queries": res["batch1_queries"], "cores": 1}}
                log(f"[bench] CPU reference hnswlib {cpu['value']} QPS on {threads} threads ({model}, nproc {nproc}); "
                    f"batch=1 p50 {res['batch1_p50_us']} us")
                return cpu
            log(f"[bench] reference CPU baseline failed ({r.returncode}): {r.stderr[-300:]}; timing the oracle")
    from oracle import oracle as orc
    og = orc.OracleGraph(g.arrays())
    probe = 200
    t0 = time.perf_counter()
    og.search(q[:probe], a.k, ef, order=orc.SEQ, nthreads=threads)
    per_q = (time.perf_counter() - t0) / probe
    reps = max(1, int(a.cpu_seconds / max(per_q * len(q), 1e-6)))
    ns = int(min(len(q), max(probe, a.cpu_seconds / max(per_q, 1e-7))))
    t0 = time.perf_counter()
    for _ in range(reps):
        ids, _, _ = og.search(q[:ns], a.k, ef, order=orc.SEQ, nthreads=threads)
    cs = time.perf_counter() - t0
    cpu = {"value": round(reps * ns / cs, 1), "unit": "QPS", "cores": threads, "kind": "port",
           "sample": f"{reps} pass(es) over the first {ns} of the {len(q)} queries ({cs:.1f} s), ef={ef}, k={a.k}, "
                     f"hnswlib-order oracle (oracle/oracle.c searchKnn), OpenMP over queries on {threads} host threads",
           "recall_at_10": round(recall_at(ids, gt[:ns], a.k), 4), "cpu_model": model, "nproc": nproc}
    lat = []
    for i in range(502):
        t0 = time.perf_counter()
        og.search(q[i:i + 1], a.k, ef, order=orc.SEQ, nthreads=1)
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat[1:-1]) * 1e6
    cpu["batch1_latency_us"] = {"p50": round(float(np.percentile(lat, 50)), 1),
                                "p99": round(float(np.percentile(lat, 99)), 1), "queries": 500, "cores": 1}
    log(f"[bench] CPU oracle {cpu['value']} QPS on {threads} threads")
    return cpu


FP64_MFMA_PEAK_TFLOPS = 78.6


def gt_only(a, gvs):
    """Time gvs_bruteforce_f64_device on nq x n x dim (the graph is a ring,
    unused by the brute force): 1 warm-up and 3 timed calls with HIP events;
    FLOP = 2 * nq * n * dim."""
    from gvs import hip as H
    x = gvs.gen_mixture(a.n, a.dim, a.centres, 1234, 1235, nthreads=a.build_threads)
    q = gvs.gen_mixture(a.nq, a.dim, a.centres, 1234, 98765, nthreads=a.build_threads)
    ring = ((np.arange(a.n, dtype=np.int64) + 1) % a.n).astype(np.int32).reshape(a.n, 1)

    class Ring:
        view = gvs.Graph()
    v = Ring.view
    v.n, v.dim, v.metric, v.vectors, v.ld_vectors = a.n, a.dim, gvs.L2, x.ctypes.data, a.dim
    v.r0, v.adj0, v.max_level, v.entry, v.ru, v.up_rows = 1, ring.ctypes.data, 0, 0, 0, 0
    dev = gvs.DeviceIndex(Ring)
    st = H.Stream()
    dq = H.DeviceArray.from_host(q)
    ids = H.DeviceArray((a.nq, a.k), np.int32)
    ds = H.DeviceArray((a.nq, a.k), np.float64)
    dev.bruteforce_f64_device(dq.address, a.nq, a.dim, a.k, ids.address, ds.address, st.handle)
    st.synchronize()
    e0, e1 = H.Event(), H.Event()
    e0.record(st)
    for _ in range(3):
        dev.bruteforce_f64_device(dq.address, a.nq, a.dim, a.k, ids.address, ds.address, st.handle)
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / 3
    tflops = 2.0 * a.nq * a.n * a.dim / (ms / 1e3) / 1e12
    print(json.dumps({"metric": "FP64 brute-force kNN (MFMA) time", "ms": round(ms, 3), "tflops": round(tflops, 2),
                      "peak_tflops": FP64_MFMA_PEAK_TFLOPS, "frac": round(tflops / FP64_MFMA_PEAK_TFLOPS, 4),
                      "config": {"workload": f"{a.nq} queries x {a.n} x {a.dim} fp32 (FP64 arithmetic), k={a.k}"}}),
          flush=True)
    del Ring, dev


def serving_leg(a, cache, ef):
    """Serving latency in a child process (bounded by a timeout, so that the
    batch figures above are reported whatever happens to the serving loop)."""
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--serve-child", "--graph-cache", cache,
           "--n", str(a.n), "--dim", str(a.dim), "--centres", str(a.centres), "--nq", str(a.nq),
           "--k", str(a.k), "--serve-ef", str(ef), "--build-threads", str(a.build_threads)]
    log(f"[bench] serving leg: {' '.join(cmd[2:])}")
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, timeout=300)
    except subprocess.TimeoutExpired:
        log("[bench] serving leg timed out")
        return None
    if r.returncode != 0:
        log(f"[bench] serving leg failed with status {r.returncode}")
        return None
    lines = [ln for ln in r.stdout.decode().splitlines() if ln.startswith("{")]
    return json.loads(lines[-1]) if lines else None


def precision_c4(a, gvs, g, x, q):
    """Traversal over fp32 / BF16 / FP8-e4m3 copies of the vectors (fp32
    accumulation), with and without an exact fp32 rerank of the final ef
    candidates; recall@10 and returned-distance error against FP64."""
    from gvs import hip as H
    ef = 128
    t0 = time.time()
    nsg = g.to_nsg(64)
    log(f"[c4] NSG-style graph (out-degree <= 64, navigating node {nsg.view.entry}) in {time.time() - t0:.1f}s")
    dev = gvs.DeviceIndex(nsg, storages=(gvs.F32, gvs.BF16, gvs.FP8))
    st = H.Stream()
    nq = len(q)
    dq = H.DeviceArray.from_host(q)
    gti = H.DeviceArray((nq, a.k), np.int32)
    gtd = H.DeviceArray((nq, a.k), np.float64)
    t0 = time.time()
    dev.bruteforce_f64_device(dq.address, nq, a.dim, a.k, gti.address, gtd.address, st.handle)
    st.synchronize()
    gt = gti.to_host()
    log(f"[c4] FP64 ground truth {time.time() - t0:.1f}s")
    dl = H.DeviceArray((nq, a.k), np.int64)
    dd = H.DeviceArray((nq, a.k), np.float32)
    rows = []
    names = {gvs.F32: "fp32", gvs.BF16: "bf16", gvs.FP8: "fp8_e4m3"}
    for storage, rerank in ((gvs.F32, False), (gvs.BF16, False), (gvs.BF16, True), (gvs.FP8, False),
                            (gvs.FP8, True)):
        kw = dict(k=a.k, ef=ef, storage=storage, rerank=rerank)
        dev.search_device(dq.address, nq, a.dim, dl.address, dd.address, stream=st.handle, **kw)
        st.synchronize()
        lab, dist = dl.to_host(), dd.to_host()
        e0, e1 = H.Event(), H.Event()
        e0.record(st)
        for _ in range(3):
            dev.search_device(dq.address, nq, a.dim, dl.address, dd.address, stream=st.handle, asynchronous=True,
                              **kw)
        e1.record(st)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / 3
        # FP64 distance of every returned id
        ok = lab >= 0
        xs = x[np.where(ok, lab, 0)].astype(np.float64)
        d64 = ((xs - q[:, None, :].astype(np.float64)) ** 2).sum(-1)
        rel = np.abs(dist.astype(np.float64) - d64) / np.maximum(np.abs(d64), 1e-12)
        rel = rel[ok]
        r = {"storage": names[storage], "rerank": rerank, "recall_at_10": round(recall_at(lab, gt, a.k), 4),
             "qps": round(nq / (ms / 1e3), 1), "ms_per_batch": round(ms, 3),
             "dist_rel_err_mean": float(f"{rel.mean():.3e}"), "dist_rel_err_max": float(f"{rel.max():.3e}")}
        # the library autotuner's launch for this storage (DST mc / mg, persistent grid, waves per query)
        cfg, _ = dev.autotune(dq.address, nq, a.dim, dl.address, dd.address, stream=st.handle, **kw)
        tkw = dict(kw, **cfg)
        dev.search_device(dq.address, nq, a.dim, dl.address, dd.address, stream=st.handle, **tkw)
        st.synchronize()
        tlab = dl.to_host()
        e0.record(st)
        for _ in range(3):
            dev.search_device(dq.address, nq, a.dim, dl.address, dd.address, stream=st.handle, asynchronous=True,
                              **tkw)
        e1.record(st)
        e1.synchronize()
        tms = e0.elapsed_ms(e1) / 3
        r["tuned"] = {"config": {kk: v for kk, v in cfg.items() if kk not in ("k", "ef")},
                      "recall_at_10": round(recall_at(tlab, gt, a.k), 4), "qps": round(nq / (tms / 1e3), 1)}
        rows.append(r)
        log(f"[c4] {r}")
    # NSG search from a pool (nsg.py:62-125): the navigating node's first
    # ef neighbours, and (fp32) the same filled up to ef with random ids as
    # nsg.py draws them (numpy's generator standing in for Python's random)
    a_arr = nsg.arrays()
    head = [int(v) for v in a_arr["adj0"][a_arr["entry"], :ef] if v >= 0]
    rng = np.random.default_rng(7)
    fill = np.empty((nq, ef), dtype=np.int32)
    for i in range(nq):
        taken, ids = set(head), list(head)
        while len(ids) < ef:
            v = int(rng.integers(0, a.n))
            if v not in taken:
                taken.add(v)
                ids.append(v)
        fill[i] = ids
    dfill = H.DeviceArray.from_host(fill)
    for storage, rerank, use_fill in ((gvs.F32, False, False), (gvs.F32, False, True), (gvs.BF16, True, False),
                                      (gvs.FP8, True, False)):
        kw = dict(k=a.k, ef=ef, storage=storage, rerank=rerank)
        pool = dict(init_ptr=dfill.address, init_len=ef, init_ld=ef) if use_fill else {}

        def go(asynchronous):
            dev.search_pool_device(dq.address, nq, a.dim, dl.address, dd.address, stream=st.handle,
                                   asynchronous=asynchronous, **pool, **kw)
        go(False)
        st.synchronize()
        lab = dl.to_host()
        e0, e1 = H.Event(), H.Event()
        e0.record(st)
        for _ in range(3):
            go(True)
        e1.record(st)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / 3
        r = {"storage": names[storage], "rerank": rerank,
             "start": "NSG pool: navigating node's neighbours" + (" + random fill" if use_fill else ""),
             "recall_at_10": round(recall_at(lab, gt, a.k), 4), "qps": round(nq / (ms / 1e3), 1),
             "ms_per_batch": round(ms, 3)}
        rows.append(r)
        log(f"[c4] {r}")
    vb = dev.device_bytes() if hasattr(dev, "device_bytes") else None
    print(json.dumps({"metric": "recall@10 and distance error vs FP64, low-precision traversal + fp32 rerank",
                      "config": {"workload": f"NSG-style flat graph from HNSW M={a.M} efC={a.efc} layer 0, "
                                             f"out-degree <= 64, navigating-node entry; {a.n}x{a.dim} fp32 L2, "
                                             f"{a.centres} centres; {nq} queries, k={a.k}, ef={ef}"},
                      "device_bytes": vb, "rows": rows}), flush=True)


def intra_latency(a, gvs, g, q):
    """Batch=1 latency: one query per launch through a captured graph, for
    waves per query x candidates per step (and a few pipelined DST settings,
    mg > 1) x ef; p50/p99 over 190 queries
    (10 warm-up), mean evaluations and recall@10 against FP64 ground truth."""
    from gvs import hip as H
    dev = gvs.DeviceIndex(g)
    st = H.Stream()
    nq = 200
    dq = H.DeviceArray((1, a.dim), np.float32)
    dl = H.DeviceArray((1, a.k), np.int64)
    dd = H.DeviceArray((1, a.k), np.float32)
    ds = H.DeviceArray((1, gvs.STATS_ROW), np.int32)
    dqa = H.DeviceArray.from_host(q[:nq])
    gti = H.DeviceArray((nq, a.k), np.int32)
    gtd = H.DeviceArray((nq, a.k), np.float64)
    dev.bruteforce_f64_device(dqa.address, nq, a.dim, a.k, gti.address, gtd.address, st.handle)
    st.synchronize()
    gt = gti.to_host()
    rows = []
    for ef in (64, 128, 256, 512):
        base = None
        configs = [(W, mc, 1) for W in (1, 2, 4, 8) for mc in (1, 2, 4, 8) if not (W == 8 and mc == 1)]
        configs += [(1, 1, 2), (1, 2, 2), (4, 4, 2), (8, 8, 2), (8, 8, 4)]
        for W, mc, mg in configs:
            sg = dev.capture(dq.address, 1, a.dim, dl.address, dd.address, k=a.k, ef=ef, mc=mc, mg=mg,
                             waves_per_query=W, stats_ptr=ds.address)
            lat, ev, hits = [], [], 0
            for i in range(nq):
                dq.upload(q[i:i + 1])
                t0 = time.perf_counter()
                sg.launch(st.handle)
                st.synchronize()
                lat.append(time.perf_counter() - t0)
                ev.append(int(ds.to_host()[0, 1]))
                hits += len(set(dl.to_host()[0].tolist()) & set(gt[i].tolist()))
            sg.close()
            lat = np.array(lat[10:]) * 1e6
            r = {"ef": ef, "waves": W, "mc": mc, "mg": mg, "p50_us": round(float(np.percentile(lat, 50)), 1),
                 "p99_us": round(float(np.percentile(lat, 99)), 1), "evals": round(float(np.mean(ev)), 1),
                 "recall_at_10": round(hits / (nq * a.k), 4)}
            if W == 1 and mc == 1 and mg == 1:
                base = r
            r["p50_speedup_vs_plain"] = round(base["p50_us"] / r["p50_us"], 2)
            rows.append(r)
            log(f"[intra] {r}")
    best = {ef: max((r for r in rows if r["ef"] == ef), key=lambda r: r["p50_speedup_vs_plain"]) for ef in (64, 128, 256, 512)}
    print(json.dumps({"metric": "batch=1 search latency, intra-query DST vs plain", "unit": "us",
                      "config": {"workload": f"HNSW search, {a.n}x{a.dim} fp32 L2, M={a.M}, efC={a.efc}, "
                                             f"one query per launch (captured graph), k={a.k}"},
                      "best_per_ef": best, "rows": rows}), flush=True)


def serve_child(a, gvs, g, q, ks=None, efs=None, metric_name="L2", max_inflight=4096, grid_waves=0, nreq=100000,
                waves_per_request=1, mc=1):
    """Open-loop Poisson load at fractions of the serving loop's saturated
    throughput; latency runs from a request's arrival to the poll that sees
    its result.  ks / efs: per-request k and ef (default: a.k, a.serve_ef).
    max_inflight: device slots, which also bound the serving grid.
    grid_waves: the serving grid (0 = resident waves, at most max_inflight).
    waves_per_request / mc: one wave per request, or one workgroup of that
    many waves running DST(mc, 1) (lower latency for heavy requests).  The
    query set is cycled to nreq requests; the host-side time split of the
    saturated run is reported with the latencies."""
    dev = gvs.DeviceIndex(g)
    qq = np.ascontiguousarray(np.tile(q, (nreq // len(q) + 1, 1))[:nreq])
    ks = np.full(nreq, a.k, dtype=np.int32) if ks is None else np.resize(ks, nreq).astype(np.int32)
    efs = np.full(nreq, a.serve_ef, dtype=np.int32) if efs is None else np.resize(efs, nreq).astype(np.int32)
    max_k, max_ef = int(ks.max()), int(max(efs.max(), ks.max()))
    srv = gvs.Server(dev, max_inflight=max_inflight, max_k=max_k, max_ef=max_ef, grid_waves=grid_waves, mc=mc,
                     waves_per_request=waves_per_request)
    # untimed warm-up: the first requests pay the serving kernel's first
    # launch and the first touches of the pinned rings
    nw = min(nreq, 5000)
    srv.loadgen(qq[:nw], 1e9, ks[:nw], efs[:nw], seed=0)
    _, _, lat, wall = srv.loadgen(qq, 1e9, ks, efs, seed=1)
    peak = nreq / wall
    host = srv.loadgen_breakdown()
    log(f"[serve] saturated {peak:.0f} QPS; host side {host}")
    loads = []
    for frac in (0.25, 0.5, 0.75, 0.9):
        _, _, lat, wall = srv.loadgen(qq, peak * frac, ks, efs, seed=2)
        loads.append({"load": frac, "offered_qps": round(peak * frac, 1), "achieved_qps": round(nreq / wall, 1),
                      "p50_us": round(float(np.percentile(lat, 50)), 1),
                      "p99_us": round(float(np.percentile(lat, 99)), 1)})
        log(f"[serve] load {frac:.2f}: p50 {loads[-1]['p50_us']} us p99 {loads[-1]['p99_us']} us")
    srv.close()
    kdesc = a.k if len(set(ks.tolist())) == 1 else sorted(set(ks.tolist()))
    edesc = a.serve_ef if len(set(efs.tolist())) == 1 else [int(efs.min()), int(efs.max())]
    print(json.dumps({"requests": nreq, "k": kdesc, "ef": edesc, "peak_qps": round(peak, 1),
                      "max_inflight": max_inflight, "grid_waves": grid_waves, "waves_per_request": waves_per_request,
                      "mc": mc, "workload": f"{a.n}x{a.dim} {metric_name}", "host_breakdown_saturated": host,
                      "loads": loads}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--M", type=int, default=32)
    ap.add_argument("--efc", type=int, default=200)
    ap.add_argument("--centres", type=int, default=256)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--streams", type=int, choices=(1, 2, 3, 4, 6), default=3,
                    help="consecutive timed steps go round this many streams (forks of the index on all but the first)")
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--target", type=float, default=0.95)
    ap.add_argument("--efs", default="16,24,32,40,48,52,56,58,60,62,64,96,128,192,256",
                    help="ef grid of the plain sweep; the autotuner tries every ef of it up to the "
                         "plain path's first ef that reaches the recall target")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--build-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", default="",
                    help="fixed search config, e.g. ef=64,mc=2,persistent=1 (skips the sweep and the autotuner)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run the library's streaming copy kernel over this many MiB (counter calibration)")
    ap.add_argument("--graph-cache", default="",
                    help="hnswlib .bin file: loaded if present, else written after the build")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-latency leg")
    ap.add_argument("--builder", choices=("cpu", "gpu"), default="cpu",
                    help="graph construction: the multithreaded CPU builder, or the GPU builder")
    ap.add_argument("--no-gpu-build", action="store_true",
                    help="skip the leg that builds the same graph on the GPU and compares its recall")
    ap.add_argument("--intra", action="store_true",
                    help="instead of the throughput run: batch=1 latency of intra-query DST "
                         "(1-8 waves per query, mc 1-8, ef 64-512) against the plain search")
    ap.add_argument("--c5", action="store_true",
                    help="instead of the throughput run: serving on 1M x 384 unit-norm inner product, "
                         "Poisson arrivals at 25/50/75/90%% of peak, k in {1,10,100}, ef uniform in [32,512]")
    ap.add_argument("--c4", action="store_true",
                    help="instead of the throughput run: NSG-style flat graph (out-degree 64, navigating-node "
                         "entry) on 10M x 128 (1024 centres), fp32 / BF16 / FP8-e4m3 traversal with and without "
                         "exact fp32 rerank, ef=128, k=10: recall and distance error against FP64")
    ap.add_argument("--serve-only", action="store_true",
                    help="instead of the throughput run: only the serving leg (saturated rate, then Poisson load "
                         "at 25/50/75/90%% of it), e.g. under a profiler; with --c5, one configuration")
    ap.add_argument("--serve-grid", type=int, default=-1, help="serving grid waves for --serve-only (-1 = default)")
    ap.add_argument("--serve-w", type=int, default=1, help="waves per request for --serve-only")
    ap.add_argument("--serve-mc", type=int, default=1, help="DST candidates per step for --serve-only")
    ap.add_argument("--serve-requests", type=int, default=100000, help="requests per load level for --serve-only")
    ap.add_argument("--gt-only", action="store_true",
                    help="instead of the throughput run: time the FP64 MFMA brute-force ground truth "
                         "(nq queries x n base vectors x dim) and report TFLOP/s against the FP64 matrix peak")
    ap.add_argument("--serve-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--serve-ef", type=int, default=64, help=argparse.SUPPRESS)
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "k_search_1Mx128_hbm_traffic.txt"),
                    help="rocprofv3 PMC summary holding HBM bytes per launch of the search kernel")
    a = ap.parse_args()

    if a.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {a.gpus}); on several, each GPU would "
              "hold a replica of the graph and vectors and take a round-robin share of the queries",
              file=sys.stderr)
        sys.exit(2)

    os.environ.setdefault("GVS_VERBOSE", "1")  # the graph builder reports its progress
    import gvs
    from gvs import hip as H

    if a.gt_only:
        gt_only(a, gvs)
        return
    t0 = time.time()
    metric = gvs.L2
    if a.c4:
        if a.n == 1_000_000:
            a.n = 10_000_000
        a.centres = 1024
    if a.c5:
        a.dim = 384
        metric = gvs.IP
    x = gvs.gen_mixture(a.n, a.dim, a.centres, 1234, 1235, normalize=a.c5, nthreads=a.build_threads)
    q = gvs.gen_mixture(a.nq, a.dim, a.centres, 1234, 98765, normalize=a.c5, nthreads=a.build_threads)
    log(f"[bench] data {a.n}x{a.dim} + {a.nq} queries in {time.time() - t0:.1f}s; building HNSW "
        f"M={a.M} efC={a.efc} on {a.build_threads} threads")
    t0 = time.time()
    if a.graph_cache and os.path.exists(a.graph_cache):
        g = gvs.HostGraph.load_hnswlib(a.graph_cache, a.dim, metric=metric)
        log(f"[bench] graph loaded from {a.graph_cache}")
    elif a.builder == "gpu":
        g, bst = gvs.HostGraph.build_hnsw_gpu(x, M=a.M, ef_construction=a.efc, metric=metric, with_stats=True)
        log(f"[bench] GPU build {bst}")
    else:
        g = gvs.HostGraph.build_hnsw(x, M=a.M, ef_construction=a.efc, metric=metric, nthreads=a.build_threads)
    build_s = time.time() - t0
    log(f"[bench] build ({a.builder}) {build_s:.1f}s")
    if a.serve_child:
        serve_child(a, gvs, g, q)
        return
    if a.c4:
        precision_c4(a, gvs, g, x, q)
        return
    if a.serve_only and not a.c5:
        if a.graph_cache and not os.path.exists(a.graph_cache):
            g.save_hnswlib(a.graph_cache)
        serve_child(a, gvs, g, q, nreq=a.serve_requests, grid_waves=max(a.serve_grid, 0),
                    waves_per_request=a.serve_w, mc=a.serve_mc)
        return
    if a.c5:
        if a.graph_cache and not os.path.exists(a.graph_cache):
            g.save_hnswlib(a.graph_cache)
        rng = np.random.default_rng(5)
        ks = rng.choice([1, 10, 100], size=len(q)).astype(np.int32)
        efs = rng.integers(32, 513, size=len(q)).astype(np.int32)
        if a.serve_only:
            serve_child(a, gvs, g, q, ks=ks, efs=efs, metric_name="inner product, unit-norm", max_inflight=4096,
                        grid_waves=max(a.serve_grid, 0), nreq=a.serve_requests, waves_per_request=a.serve_w,
                        mc=a.serve_mc)
            return
        # one wave per request (the plain search's configuration), one
        # workgroup of 4 waves running DST(8, 1) per request (the highest rate
        # and p99 under 1 ms at half load), and of 8 waves (the lowest latency)
        serve_child(a, gvs, g, q, ks=ks, efs=efs, metric_name="inner product, unit-norm", max_inflight=4096,
                    grid_waves=768, nreq=30000)
        serve_child(a, gvs, g, q, ks=ks, efs=efs, metric_name="inner product, unit-norm", max_inflight=4096,
                    grid_waves=0, nreq=30000, waves_per_request=4, mc=8)
        serve_child(a, gvs, g, q, ks=ks, efs=efs, metric_name="inner product, unit-norm", max_inflight=4096,
                    grid_waves=1024, nreq=30000, waves_per_request=8, mc=8)
        return
    if a.intra:
        intra_latency(a, gvs, g, q)
        return
    cache = a.graph_cache or f"/tmp/gvs_bench_graph_{os.getpid()}.bin"
    if not os.path.exists(cache):
        g.save_hnswlib(cache)
    dev = gvs.DeviceIndex(g)
    r0 = g.view.r0
    ru = g.view.ru

    stream = H.Stream()
    dq = H.DeviceArray.from_host(q)
    dlab = H.DeviceArray((a.nq, a.k), np.int64)
    ddist = H.DeviceArray((a.nq, a.k), np.float32)
    dstat = H.DeviceArray((a.nq, gvs.STATS_ROW), np.int32)
    gti = H.DeviceArray((a.nq, a.k), np.int32)
    gtd = H.DeviceArray((a.nq, a.k), np.float64)
    t0 = time.time()
    dev.bruteforce_f64_device(dq.address, a.nq, a.dim, a.k, gti.address, gtd.address, stream.handle)
    stream.synchronize()
    gt = gti.to_host()
    log(f"[bench] FP64 brute-force ground truth (MFMA) {time.time() - t0:.2f}s")

    def run(cfg, stats=False, asynchronous=True):
        dev.search_device(dq.address, a.nq, a.dim, dlab.address, ddist.address, k=a.k,
                          stats_ptr=dstat.address if stats else None, stream=stream.handle,
                          asynchronous=asynchronous, **cfg)

    def timed(cfg, reps):
        e0, e1 = H.Event(), H.Event()
        stream.synchronize()
        e0.record(stream)
        for _ in range(reps):
            run(cfg)
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_ms(e1) / reps

    def measure_recall(cfg):
        run(cfg, asynchronous=False)
        stream.synchronize()
        return recall_at(dlab.to_host(), gt, a.k)

    if a.copy_calib > 0:
        nb = a.copy_calib << 20
        src, dst = H.DeviceArray(nb, np.uint8), H.DeviceArray(nb, np.uint8)
        for _ in range(3):
            gvs.device_copy(dst.address, src.address, nb, stream.handle)
        stream.synchronize()
        log(f"[bench] copy calibration: 3 x {a.copy_calib} MiB through k_copy16")
        del src, dst

    # ef sweep for the plain one-wavefront-per-query search
    efs = [int(e) for e in a.efs.split(",")]
    if a.config:
        kv = dict(item.split("=") for item in a.config.split(","))
        fixed = dict(ef=int(kv.get("ef", 64)), mc=int(kv.get("mc", 1)), mg=int(kv.get("mg", 1)),
                     persistent=bool(int(kv.get("persistent", 0))))
        if int(kv.get("grid_waves", 0)):
            fixed["grid_waves"] = int(kv["grid_waves"])
        if int(kv.get("region_order", 0)):
            fixed["region_order"] = True
        if int(kv.get("waves_per_query", 1)) > 1:
            fixed["waves_per_query"] = int(kv["waves_per_query"])
        efs = [fixed["ef"]]
    plain_ef, plain_rec = None, None
    sweep = []
    for ef in efs:
        cfg = dict(ef=ef, mc=1, persistent=False)
        rec = measure_recall(cfg)
        ms = timed(cfg, 2)
        sweep.append({"ef": ef, "recall": round(rec, 4), "qps_plain": round(a.nq / (ms / 1e3))})
        log(f"[bench] plain ef={ef}: recall@10={rec:.4f} {ms:.2f} ms/batch")
        if rec >= a.target and plain_ef is None:
            plain_ef, plain_rec = ef, rec
    if plain_ef is None:
        plain_ef, plain_rec = efs[-1], rec
    plain_cfg = dict(ef=plain_ef, mc=1, persistent=False)

    # autotune the batched path with the library's autotuner (gvs_autotune:
    # mc x mg x persistent x resident waves) at each ef up to the plain one;
    # keep the fastest configuration whose recall holds the target
    best = None
    if a.config:
        best = (0.0, fixed, measure_recall(fixed))
    for ef in ([] if a.config else [e for e in efs if e <= plain_ef]):
        cfg, ms = dev.autotune(dq.address, a.nq, a.dim, dlab.address, ddist.address, k=a.k, ef=ef,
                               stream=stream.handle)
        cfg = {kk: v for kk, v in cfg.items()
               if not (kk in ("grid_waves", "region_order") and not v) and not (kk == "waves_per_query" and v == 1)}
        rec = measure_recall(cfg)
        log(f"[bench] autotune ef={ef}: {cfg} recall={rec:.4f} {ms:.3f} ms")
        if rec < a.target:
            continue
        ms = timed(cfg, 3)
        if best is None or ms < best[0]:
            best = (ms, cfg, rec)
    if best is None:
        best = (timed(plain_cfg, 2), plain_cfg, plain_rec)
    tuned_cfg, tuned_rec = best[1], best[2]
    log(f"[bench] chosen {tuned_cfg} recall={tuned_rec:.4f}")

    # Timed region: W warmup + K steps.  Consecutive steps go round the index
    # and forks of it (the same device arrays, scratch of their own,
    # gvs_index_fork), each on a stream of its own, so that one batch's
    # descent, sort and ramp run under the previous batches' drain; every
    # step is a whole search of the batch.  The clock is the host's around a
    # device-wide synchronisation on both sides (HIP events on the first
    # stream too).
    forks = [dev.fork() for _ in range(a.streams - 1)]
    fstreams = [H.Stream() for _ in forks]
    fout = [(H.DeviceArray((a.nq, a.k), np.int64), H.DeviceArray((a.nq, a.k), np.float32)) for _ in forks]
    nstreams = a.streams

    def run_step(i, cfg):
        j = i % nstreams
        if j == 0:
            run(cfg)
        else:
            forks[j - 1].search_device(dq.address, a.nq, a.dim, fout[j - 1][0].address, fout[j - 1][1].address, k=a.k,
                                       stream=fstreams[j - 1].handle, asynchronous=True, **cfg)

    def timed_steps(nsteps):
        # the warmup takes every stream at least once (a handle's first search
        # sizes its scratch)
        for i in range(max(a.warmup, nstreams)):
            run_step(i, tuned_cfg)
        H.synchronize()
        e0, e1 = H.Event(), H.Event()
        w0 = time.perf_counter()
        e0.record(stream)
        for i in range(nsteps):
            run_step(i, tuned_cfg)
        e1.record(stream)
        e1.synchronize()
        H.synchronize()
        wall = time.perf_counter() - w0
        return max(e0.elapsed_ms(e1), wall * 1e3) / nsteps

    ms_step = timed_steps(a.steps)
    if dev.overflow_count() != 0 or any(f.overflow_count() != 0 for f in forks):
        log("[bench] warning: visited-table overflow in the timed region")
    if a.steps >= a.streams and any(not np.array_equal(fl.to_host(), dlab.to_host()) for fl, _ in fout):
        raise SystemExit("a fork's results differ from the index's")
    qps = a.nq / (ms_step / 1e3)
    nstreams = 1
    ms_one_stream = timed_steps(a.steps)
    nstreams = a.streams
    plain_ms = timed(plain_cfg, a.steps)
    plain_qps = a.nq / (plain_ms / 1e3)

    # algorithmic bytes of the tuned search, from the kernel's exact counters
    run(tuned_cfg, stats=True, asynchronous=False)
    stream.synchronize()
    st = dstat.to_host().astype(np.int64)
    row_b = 4 * a.dim
    traffic = int((st[:, 0] * 4 * r0 + st[:, 1] * row_b + st[:, 2] * 4 * ru + st[:, 3] * row_b).sum())
    kernel_ms = timed(tuned_cfg, 3)
    achieved = traffic / (kernel_ms / 1e3) / 1e9
    avg = st.mean(axis=0)
    log(f"[bench] tuned {qps:.0f} QPS on {a.streams} stream(s), {a.nq / (ms_one_stream / 1e3):.0f} on one, "
        f"plain {plain_qps:.0f} QPS; hops {avg[0]:.1f} evals {avg[1]:.1f} "
        f"upper {avg[2]:.1f}/{avg[3]:.1f}; {achieved:.0f} GB/s algorithmic")

    # small batches: latency of one eager search call vs one replay of the
    # captured launch sequence (hipGraph), plain config
    small = []
    for nb in (1, 64):
        cfg_s = dict(plain_cfg)
        reps = 200

        def eager():
            dev.search_device(dq.address, nb, a.dim, dlab.address, ddist.address, k=a.k, stream=stream.handle,
                              asynchronous=True, **cfg_s)
        sg = dev.capture(dq.address, nb, a.dim, dlab.address, ddist.address, k=a.k, **cfg_s)
        res = {"batch": nb}
        for name, fn in (("eager_us", eager), ("graph_us", lambda: sg.launch(stream.handle))):
            for _ in range(10):
                fn()
            stream.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
                stream.synchronize()
            res[name] = round((time.perf_counter() - t0) / reps * 1e6, 1)
        sg.close()
        # the library's autotuner on the same small batch (mc, mg, persistent,
        # waves per query), timed as the eager call above; DST can change the
        # results, so its recall over the whole query set is reported with it
        cfg_t, _ = dev.autotune(dq.address, nb, a.dim, dlab.address, ddist.address, k=a.k, ef=plain_ef,
                                stream=stream.handle)

        def tuned_call():
            dev.search_device(dq.address, nb, a.dim, dlab.address, ddist.address, k=a.k, stream=stream.handle,
                              asynchronous=True, **cfg_t)
        for _ in range(10):
            tuned_call()
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            tuned_call()
            stream.synchronize()
        res["autotuned"] = {"us": round((time.perf_counter() - t0) / reps * 1e6, 1),
                            "config": {kk: v for kk, v in cfg_t.items() if kk != "ef"},
                            "recall_at_10": round(measure_recall(cfg_t), 4)}
        small.append(res)
        log(f"[bench] batch {nb}: eager {res['eager_us']} us, graph replay {res['graph_us']} us per call; "
            f"autotuned {res['autotuned']}")

    serving = None
    if not a.no_serving:
        serving = serving_leg(a, cache, plain_ef)
    traffic_hbm = read_traffic(a.traffic_file, tuned_cfg)
    if traffic_hbm is None:
        log(f"[bench] no PMC traffic summary for {tuned_cfg} in {a.traffic_file}")

    gpu_build = None
    if not a.no_gpu_build and a.builder == "cpu":
        # the same graph built on the GPU: build time and recall at the plain ef
        del dev
        t0 = time.time()
        g2, bst = gvs.HostGraph.build_hnsw_gpu(x, M=a.M, ef_construction=a.efc, with_stats=True)
        wall_b = time.time() - t0
        dev2 = gvs.DeviceIndex(g2)
        rec2 = recall_at(dev2.search(q, k=a.k, ef=plain_ef)[0], gt, a.k)
        gpu_build = {"seconds": round(wall_b, 2), "phases_s": {kk: round(v, 3) for kk, v in bst.items()
                                                                 if kk != "seconds"},
                     "recall_at_10_at_plain_ef": round(rec2, 4), "cpu_built_recall_at_10": round(plain_rec, 4),
                     "cpu_build_seconds": round(build_s, 1), "cpu_build_threads": a.build_threads}
        log(f"[bench] GPU build {gpu_build}")
        del dev2, g2

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(a, g, q, gt, plain_ef, cache)
    if not a.graph_cache and os.path.exists(cache):
        os.remove(cache)

    out = {
        "metric": METRIC, "value": round(qps, 1), "unit": "QPS", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic: seeded {a.centres}-centre Gaussian mixture (centres N(0,16I), noise N(0,I)); "
                "queries from the same mixture; FP64 brute-force ground truth",
        "config": {
            "workload": f"HNSW search, {a.n}x{a.dim} fp32 L2, M={a.M} (layer-0 degree {r0}), efC={a.efc}, "
                        f"{a.nq} queries per step, k={a.k}, recall@10>={a.target}",
            "ef": tuned_cfg["ef"], "mc": tuned_cfg["mc"], "mg": tuned_cfg.get("mg", 1),
            "persistent": tuned_cfg["persistent"], "grid_waves": tuned_cfg.get("grid_waves", 0),
            "region_order": bool(tuned_cfg.get("region_order", False)),
            "waves_per_query": tuned_cfg.get("waves_per_query", 1),
            "streams": a.streams, "qps_one_stream": round(a.nq / (ms_one_stream / 1e3), 1),
            "recall_at_10": round(tuned_rec, 4), "plain_ef": plain_ef, "plain_recall_at_10": round(plain_rec, 4),
            "plain_qps": round(plain_qps, 1),
            "speedup_vs_plain": round(a.nq / (ms_one_stream / 1e3) / plain_qps, 3),
            "gain_over_one_stream": round(ms_one_stream / ms_step, 3),
            "avg_hops": round(float(avg[0]), 2), "avg_dist_evals": round(float(avg[1]), 2),
            "build_seconds": round(build_s, 1), "ef_sweep": sweep, "small_batch_latency": small,
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_hbm,
                     "algorithmic_bytes_per_launch": traffic, "kernel_ms": round(kernel_ms, 4)},
        "cpu_baseline": cpu,
        "serving": serving,
        "gpu_build": gpu_build,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
