"""The no-allocation search entry point (gvs_search_device_noalloc) against
gvs_search_device: same ids, distances and counters, on two calls in a row
(the second with a smaller batch), for the LDS visited hash (ef 64), the
2-byte hash (ef 200) and the Bloom filter + HBM table, which lives in the
library's fixed scratch array."""
import numpy as np
import pytest

import gvs
from gvs import hip as H

pytestmark = pytest.mark.gpu

_G = {}


def index():
    if not _G:
        x = gvs.gen_mixture(20000, 128, 32, 71, 72)
        g = gvs.HostGraph.build_hnsw(x, M=16, ef_construction=100, nthreads=16)
        _G.update(g=g, dev=gvs.DeviceIndex(g))
    return _G["dev"]


def run_both(dev, q, k, ef, **kw):
    nq = q.shape[0]
    dq = H.DeviceArray.from_host(q)
    out = []
    for fn in (dev.search_device, dev.search_device_noalloc):
        dl = H.DeviceArray((nq, k), np.int64)
        dd = H.DeviceArray((nq, k), np.float32)
        ds = H.DeviceArray((nq, gvs.STATS_ROW), np.int32)
        fn(dq.address, nq, 128, dl.address, dd.address, k=k, ef=ef, stats_ptr=ds.address, **kw)
        H.synchronize()
        assert dev.overflow_count() == 0
        out.append((dl.to_host(), dd.to_host(), ds.to_host()[:, :gvs.STATS_PER_QUERY]))
    return out


@pytest.mark.parametrize("ef,kw", [(64, {}), (64, {"mc": 4}), (200, {}), (200, {"visited": "hbm"}),
                                   (64, {"visited": "hbm", "mc": 2})])
def test_noalloc_equals_search_device_on_two_calls(ef, kw):
    dev = index()
    for nq, seed in ((3000, 1), (257, 2)):
        q = gvs.gen_mixture(nq, 128, 32, 71, 100 + seed)
        print(f"case noalloc ef={ef} {kw} nq={nq}", flush=True)
        (a_l, a_d, a_s), (b_l, b_d, b_s) = run_both(dev, q, 10, ef, **kw)
        np.testing.assert_array_equal(b_l, a_l)
        np.testing.assert_array_equal(b_d, a_d)
        np.testing.assert_array_equal(b_s, a_s)


def test_noalloc_rejects_what_does_not_fit():
    dev = index()
    q = gvs.gen_mixture(4, 128, 32, 71, 9)
    dq = H.DeviceArray.from_host(q)
    dl = H.DeviceArray((4, 10), np.int64)
    dd = H.DeviceArray((4, 10), np.float32)
    with pytest.raises(gvs.GvsError):
        dev.search_device_noalloc(dq.address, 4, 128, dl.address, dd.address, k=10, ef=64, waves_per_query=4)
    # visited tables of 2^24 words (64 MiB each) leave no room beside the status words
    with pytest.raises(gvs.GvsError):
        dev.search_device_noalloc(dq.address, 4, 128, dl.address, dd.address, k=10, ef=64, visited="hbm",
                                  visited_log2=24)
