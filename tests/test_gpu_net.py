"""Network front end (gvs_net_serve) with the reference's wire format
(networked_FPGA/CPU_programs/README.md:40-72), driven by a client that
follows networked_FPGA/CPU_programs/CPU_client.cpp over loopback."""
import socket
import threading

import numpy as np
import pytest

import gvs
from gvs import net
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GVS_ERR_IO = -6

_G = {}


def index(dim, metric=gvs.L2):
    key = (dim, metric)
    if key not in _G:
        x = gvs.gen_mixture(12000, dim, 32, 71, 72, normalize=metric == gvs.IP)
        g = gvs.HostGraph.build_hnsw(x, M=16, ef_construction=100, metric=metric, nthreads=16)
        _G[key] = (g, gvs.DeviceIndex(g))
    return _G[key]


def session(srv, q, ef, batch_size, **kw):
    box = {}

    def start(f2c_port):
        box["node"] = net.NetNode(srv, q.shape[1], ef, f2c_port)
        return box["node"].bound_port()

    ids, dists, lat, wall = net.run_client(start, q, ef, batch_size, **kw)
    st = box["node"].join()
    assert box["node"].rc == 0, box["node"].error
    return ids, dists, st


@pytest.mark.parametrize("dim,ef,metric", [(128, 64, gvs.L2), (100, 40, gvs.L2), (96, 10, gvs.IP)])
def test_net_session_replies_equal_the_search(dim, ef, metric):
    """Every reply of a session (ragged batches, query and batch windows as
    the reference client uses them) equals the same query searched with
    k = ef on the batch path, and the oracle's search."""
    g, dev = index(dim, metric)
    q = gvs.gen_mixture(301, dim, 32, 71, 73, normalize=metric == gvs.IP)
    srv = gvs.Server(dev, max_inflight=128, max_k=ef, max_ef=ef)
    print(f"case net dim={dim} ef={ef} metric={metric}", flush=True)
    ids, dists, st = session(srv, q, ef, batch_size=17, query_window=40, batch_window=2)
    srv.close()
    assert st["batches"] == 18 and st["queries"] == 301 and st["replies"] == 301
    hdr, qb, rb = net.frame_bytes(dim, ef)
    assert st["bytes_in"] == 19 * hdr + 301 * qb and st["bytes_out"] == 301 * rb
    lab, dd = dev.search(q, k=ef, ef=ef)
    np.testing.assert_array_equal(ids, lab.astype(np.int32))
    np.testing.assert_array_equal(dists, dd)
    og = orc.OracleGraph(g.arrays())
    oi, od, _ = og.search(q[:60], ef, ef, order=orc.WAVE)
    np.testing.assert_array_equal(ids[:60], oi)
    np.testing.assert_array_equal(dists[:60], od)


def test_net_batch_of_one_at_large_ef():
    """Batches of one query with windows of one (the reference's latency
    mode) at ef = 200: the replies equal the search."""
    g, dev = index(128)
    q = gvs.gen_mixture(40, 128, 32, 71, 74)
    ef = 200
    srv = gvs.Server(dev, max_inflight=64, max_k=ef, max_ef=ef)
    ids, dists, st = session(srv, q, ef, batch_size=1, query_window=1, batch_window=1)
    srv.close()
    assert st["batches"] == 40 and st["replies"] == 40
    lab, dd = dev.search(q, k=ef, ef=ef)
    np.testing.assert_array_equal(ids, lab.astype(np.int32))
    np.testing.assert_array_equal(dists, dd)


def test_net_client_gone_mid_batch_ends_the_session_with_an_error():
    """A client that closes its query connection inside a batch: the node
    returns GVS_ERR_IO with a message instead of waiting."""
    g, dev = index(128)
    srv = gvs.Server(dev, max_inflight=64, max_k=16, max_ef=16)
    lst = socket.socket()
    lst.bind(("127.0.0.1", 0))
    lst.listen(1)
    node = net.NetNode(srv, 128, 16, lst.getsockname()[1], timeout_ms=10000)
    port = node.bound_port()
    lst.settimeout(20)
    f2c, _ = lst.accept()
    c2f = socket.create_connection(("127.0.0.1", port), timeout=20)
    hdr = np.zeros(16, dtype=np.int32)
    hdr[0] = 3
    c2f.sendall(hdr.tobytes() + np.zeros(16 * 8, np.float32).tobytes())  # one query and a half
    c2f.close()
    node.join(30)
    assert node.rc == GVS_ERR_IO
    assert "closed" in node.error
    f2c.close()
    lst.close()
    srv.close()
