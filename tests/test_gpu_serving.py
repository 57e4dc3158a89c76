"""GPU tests of the serving loop (continuous batching over the persistent
kernel): each request's result equals the same query searched alone with its
own k and ef, whatever else is in flight, and the oracle's search with that
k and ef; admission, batched submission, cancellation, rejection and the
open-loop load generator; one wave or one workgroup of waves per request;
the 1M x 384 inner-product serving shape (k in {1, 10, 100}, ef in
[32, 512]) at reduced N."""
import time

import numpy as np
import pytest

import gvs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

_G = {}


def setup():
    if "dev" not in _G:
        x = gvs.gen_mixture(20000, 128, 32, 51, 52)
        g = gvs.HostGraph.build_hnsw(x, M=16, ef_construction=100, nthreads=16)
        _G.update(x=x, g=g, dev=gvs.DeviceIndex(g), q=gvs.gen_mixture(600, 128, 32, 51, 53),
                  og=orc.OracleGraph(g.arrays()))
    return _G


def drain(srv, n, timeout=60.0, with_stats=False):
    got = {}
    t0 = time.time()
    while len(got) < n:
        out = srv.poll(256, with_stats=with_stats)
        for j, r in enumerate(out[0]):
            got[int(r)] = tuple(o[j] for o in out[1:])
        assert time.time() - t0 < timeout, f"serving loop did not drain: {len(got)} of {n} done"
    return got


def check_against_oracle(og, q, ks, efs, labels, dists, mc=1):
    """Every served row equals the oracle's search with that request's k and
    ef (the kernel's summation order, DST(mc, 1)); -1 / +inf beyond k."""
    for k in np.unique(ks):
        for ef in np.unique(efs[ks == k]):
            sel = np.nonzero((ks == k) & (efs == ef))[0]
            ids, ds, _ = og.search(q[sel], int(k), int(ef), mc=mc, order=orc.WAVE)
            np.testing.assert_array_equal(labels[sel, :k], ids.astype(np.int64))
            np.testing.assert_array_equal(dists[sel, :k], ds)
            assert np.all(labels[sel, k:] == -1) and np.all(np.isinf(dists[sel, k:]))


def test_mixed_requests_equal_alone():
    s = setup()
    dev, q = s["dev"], s["q"]
    rng = np.random.default_rng(0)
    ks = rng.choice([1, 10, 100], size=len(q))
    efs = rng.integers(32, 257, size=len(q))
    srv = gvs.Server(dev, max_inflight=128, max_k=100, max_ef=256)
    ids = []
    pending = list(range(len(q)))
    got = {}
    t0 = time.time()
    while pending or len(got) < len(q):
        while pending:
            i = pending[0]
            r = srv.try_submit(q[i], int(ks[i]), int(efs[i]))
            if r is None:
                break
            ids.append((r, i))
            pending.pop(0)
        out = srv.poll(256, with_stats=True)
        for j, r in enumerate(out[0]):
            got[int(r)] = (out[1][j], out[2][j], out[4][j])
        assert time.time() - t0 < 60, f"serving loop stalled: {len(got)} of {len(q)} done, {len(pending)} pending"
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(q) and st["submitted"] == len(q)
    L = np.full((len(q), 100), -1, dtype=np.int64)
    D = np.full((len(q), 100), np.inf, dtype=np.float32)
    for r, i in ids:
        lab, dist, cnt = got[r]
        k, ef = int(ks[i]), int(efs[i])
        alone, adist, ast = dev.search(q[i:i + 1], k=k, ef=ef, with_stats=True)
        np.testing.assert_array_equal(lab[:k], alone[0])
        np.testing.assert_array_equal(dist[:k], adist[0])
        np.testing.assert_array_equal(cnt, ast[0])
        assert np.all(lab[k:] == -1)
        L[i], D[i] = lab, dist
    check_against_oracle(s["og"], q, ks, efs, L, D)


def test_reject_cancel_and_counters():
    s = setup()
    dev, q = s["dev"], s["q"]
    srv = gvs.Server(dev, max_inflight=8, max_k=10, max_ef=64)
    with pytest.raises(gvs.GvsError):
        srv.submit(q[0], 11, 32)       # k beyond the admitted range
    with pytest.raises(gvs.GvsError):
        srv.submit(q[0], 10, 65)       # ef beyond the admitted range
    r0 = srv.submit(q[0], 10, 64)
    r1 = srv.submit(q[1], 10, 64)
    assert srv.cancel(r1)
    got = {}
    t0 = time.time()
    while r0 not in got:
        out = srv.poll(16)
        for j, r in enumerate(out[0]):
            got[int(r)] = out[1][j]
        assert time.time() - t0 < 30
    time.sleep(0.05)
    out = srv.poll(16)
    assert r1 not in set(int(r) for r in out[0])
    st = srv.stats()
    assert st["rejected"] == 2 and st["cancelled"] == 1
    # all 8 slots can be filled; a 9th waits for a free slot
    rs = [srv.submit(q[2 + i], 10, 64) for i in range(8)] if True else []
    drain(srv, 8)
    srv.close()


def test_loadgen_poisson_matches_alone():
    s = setup()
    dev, q = s["dev"], s["q"]
    rng = np.random.default_rng(1)
    ks = rng.choice([1, 10, 100], size=len(q)).astype(np.int32)
    efs = rng.integers(32, 257, size=len(q)).astype(np.int32)
    srv = gvs.Server(dev, max_inflight=512, max_k=100, max_ef=256)
    lab, dist, lat, wall = srv.loadgen(q, 20000.0, ks, efs, seed=3)
    srv.close()
    assert np.all(lat > 0)
    print(f"loadgen: {len(q)} requests in {wall:.3f}s, p50 {np.percentile(lat, 50):.1f} us, "
          f"p99 {np.percentile(lat, 99):.1f} us", flush=True)
    for i in range(len(q)):
        k, ef = int(ks[i]), int(efs[i])
        alone, adist = dev.search(q[i:i + 1], k=k, ef=ef)
        np.testing.assert_array_equal(lab[i, :k], alone[0])
        np.testing.assert_array_equal(dist[i, :k], adist[0])
    check_against_oracle(s["og"], q, ks, efs, lab, dist)


@pytest.mark.parametrize("threads", [(3, 2), (4, 4)])
def test_loadgen_several_threads_matches_alone(threads):
    """Several submit and poll threads at once (per-ring ordered publication,
    completion rings per request ring): every request still gets exactly its
    own result, saturated and under load, with nothing lost or duplicated."""
    s = setup()
    dev, q = s["dev"], s["q"]
    rng = np.random.default_rng(11)
    nq = 4 * len(q)
    qq = np.ascontiguousarray(np.tile(q, (4, 1)))
    ks = rng.choice([1, 10, 100], size=nq).astype(np.int32)
    efs = rng.integers(32, 129, size=nq).astype(np.int32)
    srv = gvs.Server(dev, max_inflight=256, max_k=100, max_ef=128)
    S, P = threads
    for rate in (1e9, 30000.0):
        lab, dist, lat, wall = srv.loadgen(qq, rate, ks, efs, seed=5, submit_threads=S, poll_threads=P)
        assert np.all(lat > 0)
        for i in range(nq):
            k, ef = int(ks[i]), int(efs[i])
            alone, adist = dev.search(qq[i:i + 1], k=k, ef=ef)
            np.testing.assert_array_equal(lab[i, :k], alone[0])
            np.testing.assert_array_equal(dist[i, :k], adist[0])
    st = srv.stats()
    srv.close()
    assert st["completed"] == 2 * nq and st["submitted"] == 2 * nq


@pytest.mark.parametrize("storage,rerank", [(gvs.BF16, True), (gvs.FP8, False), (gvs.FP8, True)])
def test_low_precision_serving_matches_alone(storage, rerank):
    """Serving over BF16 / FP8 storage (fp32 accumulation), with or without
    the exact fp32 rerank of the final ef candidates: each request's result
    equals the same query searched alone on the batch path with that
    storage, rerank, k and ef."""
    s = setup()
    if "dev_lp" not in s:
        s["dev_lp"] = gvs.DeviceIndex(s["g"], storages=(gvs.F32, gvs.BF16, gvs.FP8))
    dev, q = s["dev_lp"], s["q"]
    rng = np.random.default_rng(21)
    ks = rng.choice([1, 10, 100], size=len(q)).astype(np.int32)
    efs = rng.integers(32, 257, size=len(q)).astype(np.int32)
    srv = gvs.Server(dev, max_inflight=256, max_k=100, max_ef=256, storage=storage, rerank=rerank)
    lab, dist, lat, wall = srv.loadgen(q, 20000.0, ks, efs, seed=7)
    srv.close()
    for i in range(len(q)):
        k, ef = int(ks[i]), int(efs[i])
        alone, adist = dev.search(q[i:i + 1], k=k, ef=ef, storage=storage, rerank=rerank)
        np.testing.assert_array_equal(lab[i, :k], alone[0])
        np.testing.assert_array_equal(dist[i, :k], adist[0])


def test_admission_by_queue_memory():
    """queue_bytes bounds the memory the in-flight requests hold (query, k
    result rows, counters, control words): three k=100 requests fill a
    budget of three, and the same budget then admits eight k=1 requests;
    memory comes back as completions are polled, and every result equals the
    query searched alone."""
    s = setup()
    dev, q = s["dev"], s["q"]
    rb = lambda k: 4 * 128 + 12 * k + 4 * gvs.STATS_ROW + 16
    with pytest.raises(gvs.GvsError):
        gvs.Server(dev, max_inflight=64, max_k=100, max_ef=128, queue_bytes=rb(100) - 1)
    budget = 3 * rb(100)
    srv = gvs.Server(dev, max_inflight=64, max_k=100, max_ef=128, queue_bytes=budget)
    big = [srv.try_submit(q[i], 100, 64) for i in range(4)]
    assert all(r is not None for r in big[:3]) and big[3] is None
    assert srv.try_submit(q[4], 1, 64) is None  # the memory is taken until the completions are polled
    got = drain(srv, 3)
    for j, i in enumerate(range(3)):
        alone, _ = dev.search(q[i:i + 1], k=100, ef=64)
        np.testing.assert_array_equal(got[big[j]][0][:100], alone[0])
    n_small = budget // rb(1)
    small = [srv.try_submit(q[10 + i], 1, 64) for i in range(n_small + 1)]
    assert all(r is not None for r in small[:n_small]) and small[n_small] is None
    got = drain(srv, n_small)
    for j in range(n_small):
        alone, _ = dev.search(q[10 + j:11 + j], k=1, ef=64)
        assert got[small[j]][0][0] == alone[0][0]
    srv.close()


def test_submit_batch_admits_a_prefix_and_publishes_it():
    """gvs_server_submit_batch admits requests in order up to the free slots;
    the rest are submitted after completions; every result equals the
    oracle's; counters add up."""
    s = setup()
    dev, q = s["dev"], s["q"][:300]
    ks = np.full(300, 10, np.int32)
    efs = np.full(300, 64, np.int32)
    srv = gvs.Server(dev, max_inflight=64, max_k=10, max_ef=64)
    ids = srv.submit_batch(q, ks, efs)
    assert 0 < len(ids) <= 64
    idx = {int(r): i for i, r in enumerate(ids)}
    nxt = len(ids)
    got = {}
    t0 = time.time()
    while len(got) < 300:
        if nxt < 300:
            more = srv.submit_batch(q[nxt:], ks[nxt:], efs[nxt:])
            for j, r in enumerate(more):
                idx[int(r)] = nxt + j
            nxt += len(more)
        out = srv.poll(256, with_status=True)
        for j, r in enumerate(out[0]):
            assert out[4][j] == 0
            got[idx[int(r)]] = (out[1][j], out[2][j])
        assert time.time() - t0 < 60, f"{len(got)} of 300 done"
    st = srv.stats()
    srv.close()
    assert st["submitted"] == 300 and st["completed"] == 300 and st["overflowed"] == 0
    L = np.stack([got[i][0] for i in range(300)])
    D = np.stack([got[i][1] for i in range(300)])
    check_against_oracle(s["og"], q, ks, efs, L, D)
    with pytest.raises(gvs.GvsError):
        srv2 = gvs.Server(dev, max_inflight=8, max_k=10, max_ef=64)
        try:
            srv2.submit_batch(q[:2], np.array([10, 10], np.int32), np.array([65, 64], np.int32))
        finally:
            srv2.close()


@pytest.mark.parametrize("W", [2, 4])
def test_workgroup_per_request_server(W):
    """waves_per_request = W: one workgroup of W waves per request (DST(mc, 1)
    with mc = W): results equal the batch path's intra-query search alone
    and the oracle's DST(mc, 1)."""
    s = setup()
    dev, q = s["dev"], s["q"]
    rng = np.random.default_rng(7)
    ks = rng.choice([1, 10, 100], size=len(q)).astype(np.int32)
    efs = rng.integers(32, 257, size=len(q)).astype(np.int32)
    srv = gvs.Server(dev, max_inflight=256, max_k=100, max_ef=256, mc=W, waves_per_request=W)
    print(f"case serving W={W}", flush=True)
    lab, dist, lat, wall = srv.loadgen(q, 20000.0, ks, efs, seed=4)
    srv.close()
    for i in range(0, len(q), 7):
        k, ef = int(ks[i]), int(efs[i])
        alone, adist = dev.search(q[i:i + 1], k=k, ef=ef, mc=W, waves_per_query=W)
        np.testing.assert_array_equal(lab[i, :k], alone[0])
        np.testing.assert_array_equal(dist[i, :k], adist[0])
    check_against_oracle(s["og"], q, ks, efs, lab, dist, mc=W)


@pytest.mark.parametrize("W", [1, 4])
def test_inner_product_d384_mixed_k_and_ef(W):
    """The 1M x 384 unit-norm inner-product serving shape at 30k points:
    k in {1, 10, 100}, ef uniform in [32, 512], a server admitting up to
    ef 512; every result equals the oracle's and the query searched alone."""
    if "ip" not in _G:
        x = gvs.gen_mixture(30000, 384, 64, 57, 58, normalize=True)
        g = gvs.HostGraph.build_hnsw(x, M=32, ef_construction=200, metric=gvs.IP, nthreads=16)
        _G["ip"] = dict(dev=gvs.DeviceIndex(g), og=orc.OracleGraph(g.arrays()),
                        q=gvs.gen_mixture(400, 384, 64, 57, 59, normalize=True))
    c = _G["ip"]
    rng = np.random.default_rng(11)
    ks = rng.choice([1, 10, 100], size=400).astype(np.int32)
    efs = rng.integers(32, 513, size=400).astype(np.int32)
    mc = 1 if W == 1 else 4
    srv = gvs.Server(c["dev"], max_inflight=512, max_k=100, max_ef=512, mc=mc, waves_per_request=W)
    print(f"case serving IP d384 W={W}", flush=True)
    lab, dist, lat, wall = srv.loadgen(c["q"], 5000.0, ks, efs, seed=6)
    srv.close()
    check_against_oracle(c["og"], c["q"], ks, efs, lab, dist, mc=mc)
    for i in range(0, 400, 13):
        k, ef = int(ks[i]), int(efs[i])
        alone, adist = c["dev"].search(c["q"][i:i + 1], k=k, ef=ef, mc=mc, waves_per_query=W)
        np.testing.assert_array_equal(lab[i, :k], alone[0])
        np.testing.assert_array_equal(dist[i, :k], adist[0])
