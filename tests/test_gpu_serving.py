"""GPU tests of the serving loop (continuous batching over the persistent
kernel): each request's result equals the same query searched alone with its
own k and ef, whatever else is in flight; admission, cancellation, rejection
and the open-loop load generator."""
import time

import numpy as np
import pytest

import gvs

pytestmark = pytest.mark.gpu

_G = {}


def setup():
    if "dev" not in _G:
        x = gvs.gen_mixture(20000, 128, 32, 51, 52)
        g = gvs.HostGraph.build_hnsw(x, M=16, ef_construction=100, nthreads=16)
        _G.update(x=x, g=g, dev=gvs.DeviceIndex(g), q=gvs.gen_mixture(600, 128, 32, 51, 53))
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
    for r, i in ids:
        lab, dist, cnt = got[r]
        k, ef = int(ks[i]), int(efs[i])
        alone, adist, ast = dev.search(q[i:i + 1], k=k, ef=ef, with_stats=True)
        np.testing.assert_array_equal(lab[:k], alone[0])
        np.testing.assert_array_equal(dist[:k], adist[0])
        np.testing.assert_array_equal(cnt, ast[0])
        assert np.all(lab[k:] == -1)


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
