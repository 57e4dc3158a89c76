"""Network front end (gvs_net_serve) and a client with the reference's wire
format (networked_FPGA/CPU_programs/README.md:40-72).

The client follows networked_FPGA/CPU_programs/CPU_client.cpp: it listens
for the node's result connection first (thread_F2C, :572-619), then connects
to the node's query port (thread_C2F, :490-545) and sends each batch as a
64-byte header (int32 batch size) followed by the queries in 64-byte
packets, at most `query_window` queries ahead of the replies it has read and
at most `batch_window` batches ahead (:68-130, :180-181); a header of -1 ends the
session.  Replies come one per query, in order: header (int32 ef), ef int32
ids, ef fp32 distances, each part padded to 64 bytes.
"""
import ctypes as C
import socket
import threading
import time

import numpy as np

from . import NetConfig, _check, lib

PACKET = 64


def frame_bytes(dim, ef):
    """(batch header, one query, one reply) sizes in bytes."""
    out = (C.c_int64 * 3)()
    _check(lib.gvs_net_frame_bytes(dim, ef, out))
    return tuple(out)


class NetNode:
    """gvs_net_serve on an open Server, run in a thread (the call blocks until
    the client ends the session)."""

    def __init__(self, server, dim, ef, f2c_port, f2c_host="127.0.0.1", c2f_port=0, timeout_ms=20000):
        self._host = f2c_host.encode()
        self.cfg = NetConfig(c2f_port, self._host, f2c_port, dim, ef, timeout_ms)
        self.server = server
        self.port = C.c_int32(0)
        self.stats = (C.c_int64 * 6)()
        self.rc = None
        self.error = None
        self._t = threading.Thread(target=self._run, daemon=True)
        self._t.start()

    def _run(self):
        self.rc = lib.gvs_net_serve(self.server._h, C.byref(self.cfg), C.byref(self.port), self.stats)
        if self.rc != 0:
            self.error = lib.gvs_last_error().decode()

    def bound_port(self, timeout=20.0):
        t0 = time.time()
        while self.port.value == 0:
            if self.rc is not None or time.time() - t0 > timeout:
                raise RuntimeError(f"node did not start listening: {self.error}")
            time.sleep(0.001)
        return self.port.value

    def join(self, timeout=60.0):
        self._t.join(timeout)
        if self._t.is_alive():
            raise RuntimeError("node session did not end")
        return dict(zip(("batches", "queries", "replies", "reruns", "bytes_in", "bytes_out"), list(self.stats)))


def _recv_exact(sock, n):
    buf = bytearray(n)
    view = memoryview(buf)
    got = 0
    while got < n:
        r = sock.recv_into(view[got:], n - got)
        if r == 0:
            raise ConnectionError("connection closed")
        got += r
    return bytes(buf)


def run_client(start_node, queries, ef, batch_size, query_window=None, batch_window=None, host="127.0.0.1"):
    """Send `queries` in batches to a node and read the replies.

    start_node(f2c_port) must start the node (it connects back to that port)
    and return the node's query port.  Returns (ids [nq, ef] int32, dists
    [nq, ef] float32, per-batch latency in seconds, wall seconds)."""
    q = np.ascontiguousarray(queries, dtype=np.float32)
    nq, dim = q.shape
    hdr_b, q_b, r_b = frame_bytes(dim, ef)
    wire = np.zeros((nq, q_b // 4), dtype=np.float32)
    wire[:, :dim] = q
    nb = (nq + batch_size - 1) // batch_size
    query_window = query_window or nq
    batch_window = batch_window or nb
    lst = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
    lst.setsockopt(socket.SOL_SOCKET, socket.SO_REUSEADDR, 1)
    lst.bind((host, 0))
    lst.listen(1)
    c2f_port = start_node(lst.getsockname()[1])
    lst.settimeout(30)
    f2c, _ = lst.accept()
    lst.close()
    c2f = socket.create_connection((host, c2f_port), timeout=30)
    for s in (c2f, f2c):
        s.setsockopt(socket.IPPROTO_TCP, socket.TCP_NODELAY, 1)
    qslots = threading.Semaphore(query_window)
    bslots = threading.Semaphore(batch_window)
    t_start = [0.0] * nb
    t_end = [0.0] * nb
    ids = np.zeros((nq, ef), dtype=np.int32)
    dists = np.zeros((nq, ef), dtype=np.float32)
    err = []

    def receiver():
        try:
            for b in range(nb):
                for i in range(b * batch_size, min(nq, (b + 1) * batch_size)):
                    r = _recv_exact(f2c, r_b)
                    hdr = np.frombuffer(r[:4], dtype=np.int32)[0]
                    if hdr != ef:
                        raise ValueError(f"reply header {hdr}, expected ef {ef}")
                    ids[i] = np.frombuffer(r, dtype=np.int32, count=ef, offset=PACKET)
                    dists[i] = np.frombuffer(r, dtype=np.float32, count=ef, offset=(r_b - PACKET) // 2 + PACKET)
                    qslots.release()
                t_end[b] = time.perf_counter()
                bslots.release()
        except Exception as e:  # reported to the caller
            err.append(e)

    t0 = time.perf_counter()
    rt = threading.Thread(target=receiver, daemon=True)
    rt.start()
    header = np.zeros(hdr_b // 4, dtype=np.int32)
    for b in range(nb):
        bslots.acquire()
        lo, hi = b * batch_size, min(nq, (b + 1) * batch_size)
        t_start[b] = time.perf_counter()
        header[0] = hi - lo
        c2f.sendall(header.tobytes())
        for i in range(lo, hi):
            qslots.acquire()
            c2f.sendall(wire[i].tobytes())
    header[0] = -1  # end of session, sent as soon as the last batch is out (CPU_client.cpp:536-539)
    c2f.sendall(header.tobytes())
    rt.join(120)
    wall = time.perf_counter() - t0
    c2f.close()
    f2c.close()
    if err:
        raise err[0]
    return ids, dists, np.array(t_end) - np.array(t_start), wall
