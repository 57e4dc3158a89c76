"""GPU parity tests: the HIP search (through the C ABI) against the oracle.

Bar: ids, distances and the per-query counters (hops, evaluations, upper
hops, upper evaluations) equal to the oracle's in the kernel's summation
order (bit-exact), on the same seeded inputs.  Against hnswlib's scalar order
the distances agree within 1e-5 relative (tests/test_oracle.py)."""
import numpy as np
import pytest

import gvs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

_CACHE = {}


def graph(n, dim, M, efc, centres=32, metric=gvs.L2, seed=1, normalize=False):
    key = (n, dim, M, efc, centres, metric, seed, normalize)
    if key not in _CACHE:
        x = gvs.gen_mixture(n, dim, centres, seed, seed + 1, normalize=normalize)
        g = gvs.HostGraph.build_hnsw(x, M=M, ef_construction=efc, metric=metric, nthreads=16)
        _CACHE[key] = (x, g, gvs.DeviceIndex(g))
    return _CACHE[key]


def queries(nq, dim, centres=32, seed=1, normalize=False):
    return gvs.gen_mixture(nq, dim, centres, seed, seed + 99, normalize=normalize)


def check_equal(dev, og, q, k, ef, mc=1, mg=1, entries=None, **kw):
    lab, dist, st = dev.search(q, k=k, ef=ef, mc=mc, mg=mg, entries=entries, with_stats=True, **kw)
    ids, ds, ost = og.search(q, k, ef, mc=mc, mg=mg, order=orc.WAVE, entries=entries)
    np.testing.assert_array_equal(lab, ids.astype(np.int64))
    np.testing.assert_array_equal(dist, ds)
    np.testing.assert_array_equal(st.astype(np.int64), ost)
    return lab, dist, st


@pytest.mark.parametrize("ef", [10, 64, 200])
def test_plain_l2_d128_matches_oracle(ef):
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(300, 128)
    print(f"case plain L2 d128 ef={ef}", flush=True)
    check_equal(dev, og, q, 10, ef)


@pytest.mark.parametrize("mc", [2, 4, 8])
def test_dst_multi_candidate_matches_oracle(mc):
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(300, 128, seed=3)
    print(f"case dst mc={mc}", flush=True)
    check_equal(dev, og, q, 10, 64, mc=mc)
    # pools of several 64-entry chunks, merged with up to mc * 64 new keys
    check_equal(dev, og, q, 10, 200, mc=mc)
    check_equal(dev, og, q, 100, 400, mc=mc, persistent=True, region_order=True)


@pytest.mark.parametrize("mc,mg", [(1, 2), (2, 2), (4, 2), (1, 3), (2, 4), (4, 8), (8, 4)])
def test_dst_pipelined_groups_match_oracle(mc, mg):
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(300, 128, seed=14)
    print(f"case dst mc={mc} mg={mg}", flush=True)
    check_equal(dev, og, q, 10, 64, mc=mc, mg=mg)
    check_equal(dev, og, q, 10, 200, mc=mc, mg=mg)


def test_persistent_equals_plain():
    x, g, dev = graph(20000, 128, 16, 100)
    q = queries(2000, 128, seed=4)
    a = dev.search(q, k=10, ef=64, with_stats=True)
    b = dev.search(q, k=10, ef=64, with_stats=True, persistent=True)
    c = dev.search(q, k=10, ef=64, with_stats=True, persistent=True, grid_waves=64)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
    for u, v in zip(a, c):
        np.testing.assert_array_equal(u, v)


@pytest.mark.parametrize("dim,M", [(96, 16), (384, 12)])
def test_ip_unit_norm_matches_oracle(dim, M):
    x, g, dev = graph(8000, dim, M, 80, metric=gvs.IP, normalize=True)
    og = orc.OracleGraph(g.arrays())
    q = queries(200, dim, normalize=True)
    print(f"case IP d{dim}", flush=True)
    check_equal(dev, og, q, 10, 64)
    check_equal(dev, og, q, 100, 128)


def test_entries_skip_descent():
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(200, 128, seed=5)
    ent = np.random.default_rng(0).integers(0, 20000, size=200).astype(np.int32)
    lab, dist, st = check_equal(dev, og, q, 10, 64, entries=ent)
    assert np.all(st[:, 2] == 0) and np.all(st[:, 3] == 0)


def test_edge_cases():
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(64, 128, seed=6)
    check_equal(dev, og, q, 1, 1)            # ef = k = 1
    check_equal(dev, og, q, 50, 10)          # k > ef -> ef_eff = k
    check_equal(dev, og, q, 100, 256)
    lab, dist = dev.search(q[:0], k=10, ef=64)
    assert lab.shape == (0, 10)


def test_duplicate_vectors_tie_break_by_id():
    base = gvs.gen_mixture(3000, 32, 8, 7, 8)
    x = np.concatenate([base, base[:1000]])  # exact duplicates -> equal distances
    g = gvs.HostGraph.build_hnsw(x, M=8, ef_construction=64, nthreads=16)
    dev = gvs.DeviceIndex(g)
    og = orc.OracleGraph(g.arrays())
    q = np.concatenate([x[:50], queries(50, 32, centres=8, seed=7)])
    check_equal(dev, og, q, 10, 40)


def test_tiny_graphs():
    for n in (1, 2, 5):
        x = gvs.gen_mixture(n, 32, 2, 9, 10)
        g = gvs.HostGraph.build_hnsw(x, M=4, ef_construction=16, nthreads=1)
        dev = gvs.DeviceIndex(g)
        og = orc.OracleGraph(g.arrays())
        q = queries(8, 32, centres=2, seed=9)
        check_equal(dev, og, q, 3, 10)


def test_flat_nsg_style_graph():
    x, g, dev = graph(20000, 128, 16, 100)
    flat = g.to_nsg(32)
    fdev = gvs.DeviceIndex(flat)
    og = orc.OracleGraph(flat.arrays())
    q = queries(200, 128, seed=11)
    check_equal(fdev, og, q, 10, 64)


def _nsg_pools(nsg, nq, L, seed):
    """nsg.py:66-80's start: the navigating node's first neighbours, then
    distinct random ids not yet taken (numpy stands in for Python's
    random.randint) up to L."""
    a = nsg.arrays()
    head = [int(v) for v in a["adj0"][a["entry"], :min(L, a["r0"])] if v >= 0]
    rng = np.random.default_rng(seed)
    out = np.empty((nq, L), dtype=np.int32)
    for i in range(nq):
        taken, ids = set(head), list(head)
        while len(ids) < L:
            v = int(rng.integers(0, a["n"]))
            if v not in taken:
                taken.add(v)
                ids.append(v)
        out[i] = ids
    return out


@pytest.mark.parametrize("ef", [32, 100, 200])
def test_nsg_pool_search_matches_oracle(ef):
    """The NSG search (nsg.py:62-125) from the navigating node's neighbours,
    and from nsg.py's pool with random fill, equals the oracle's pool search
    and its separate restatement of nsg.py (ids, distances, expansions,
    evaluations); with DST on top it equals the oracle's DST from the pool."""
    x, g, dev = graph(20000, 128, 16, 100)
    nsg = g.to_nsg(64)
    ndev = gvs.DeviceIndex(nsg)
    og = orc.OracleGraph(nsg.arrays())
    q = queries(120, 128, seed=13)
    a = nsg.arrays()
    head = a["adj0"][a["entry"], :min(ef, a["r0"])]
    print(f"case nsg pool ef={ef} default start", flush=True)
    lab, dist, st = ndev.search_pool(q, k=10, ef=ef, with_stats=True)
    ids, ds, ost = orc.search_pool(og, q, 10, ef, head)
    np.testing.assert_array_equal(lab, ids)
    np.testing.assert_array_equal(dist, ds)
    np.testing.assert_array_equal(st[:, :2].astype(np.int64), ost)
    assert np.all(st[:, 2:] == 0)  # no descent
    pools = _nsg_pools(nsg, 120, ef, seed=ef)
    print(f"case nsg pool ef={ef} random fill", flush=True)
    lab, dist, st = ndev.search_pool(q, k=10, ef=ef, init=pools, with_stats=True)
    for ids, ds, ost in (orc.search_pool(og, q, 10, ef, pools), orc.nsg_search(og, q, 10, ef, pools)):
        np.testing.assert_array_equal(lab, ids)
        np.testing.assert_array_equal(dist, ds)
        np.testing.assert_array_equal(st[:, :2].astype(np.int64), ost)
    for mc, mg, kw in ((4, 1, {}), (2, 2, {}), (1, 1, dict(persistent=True)), (1, 1, dict(visited="hbm"))):
        print(f"case nsg pool ef={ef} mc={mc} mg={mg} {kw}", flush=True)
        lab, dist, st = ndev.search_pool(q, k=10, ef=ef, init=pools, mc=mc, mg=mg, with_stats=True, **kw)
        ids, ds, ost = orc.search_pool(og, q, 10, ef, pools, mc=mc, mg=mg)
        np.testing.assert_array_equal(lab, ids)
        np.testing.assert_array_equal(dist, ds)
        np.testing.assert_array_equal(st[:, :2].astype(np.int64), ost)


def test_pool_search_repeats_out_of_range_and_errors():
    """Pools with repeated ids, -1 and ids past n (skipped), a pool of one
    (the entry: the plain search), and the calls it refuses."""
    x, g, dev = graph(20000, 128, 16, 100)
    nsg = g.to_nsg(64)
    ndev = gvs.DeviceIndex(nsg)
    og = orc.OracleGraph(nsg.arrays())
    q = queries(60, 128, seed=14)
    rng = np.random.default_rng(9)
    pools = rng.integers(-3, 20010, size=(60, 90)).astype(np.int32)
    pools[:, 5] = pools[:, 0]
    pools[:, 70] = pools[:, 2]
    for ef in (90, 200):
        lab, dist, st = ndev.search_pool(q, k=10, ef=ef, init=pools, with_stats=True)
        ids, ds, ost = orc.search_pool(og, q, 10, ef, pools)
        np.testing.assert_array_equal(lab, ids)
        np.testing.assert_array_equal(dist, ds)
        np.testing.assert_array_equal(st[:, :2].astype(np.int64), ost)
    ent = np.full(60, nsg.view.entry, dtype=np.int32)
    lab, dist = ndev.search_pool(q, k=10, ef=64, init=ent[:, None])
    lab2, dist2 = ndev.search(q, k=10, ef=64, entries=ent)
    np.testing.assert_array_equal(lab, lab2)
    np.testing.assert_array_equal(dist, dist2)
    with pytest.raises(gvs.GvsError):
        ndev.search_pool(q, k=10, ef=64, init=pools)  # 90 ids > ef
    with pytest.raises(gvs.GvsError):
        ndev.search_pool(q, k=10, ef=64, waves_per_query=2)


def test_visited_table_overflow_retry_is_exact():
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(100, 128, seed=12)
    # smallest table: ef=400 overflows it, the library retries with larger tables;
    # at ef=1000 the first retry (2^13 slots) overflows again and the second runs
    check_equal(dev, og, q, 10, 400, visited_log2=11, visited="hbm")
    check_equal(dev, og, q, 10, 1000, visited_log2=11, visited="hbm")


@pytest.mark.parametrize("mc", [1, 4])
def test_large_ef_lds_hash_matches_hbm_table_and_oracle(mc):
    """ef in (128, 512]: the default exact LDS hash (for a batch of 100 queries
    far below the resident grid: 2^13 4-byte slots, whose 32 KiB cost no
    usable waves) gives the oracle's results and the same as the HBM table
    behind a Bloom filter.  Forced to 2-byte slots of 2^10 at ef=200, every
    query outgrows the hash, freezes it and goes on with the HBM table
    (frozen 2-byte lookups), still exact.  ef=1000 uses the Bloom filter and
    HBM table by default."""
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(100, 128, seed=18)
    for ef in (200, 400):
        plan = dev.plan(100, k=10, ef=ef, mc=mc)
        assert (plan["lds_hash_log2"], plan["lds_hash_2byte"]) == (13, 0), plan
        print(f"case large ef mc={mc} ef={ef}", flush=True)
        a = check_equal(dev, og, q, 10, ef, mc=mc)
        b = dev.search(q, k=10, ef=ef, mc=mc, with_stats=True, visited="hbm")
        for u, v in zip(a, b):
            np.testing.assert_array_equal(u, v)
    print(f"case large ef mc={mc} ef=200 2-byte hash of 2^10 slots (freeze)", flush=True)
    assert dev.plan(100, k=10, ef=200, mc=mc, visited="lds", visited_log2=10)["lds_hash_2byte"] == 1
    check_equal(dev, og, q, 10, 200, mc=mc, visited="lds", visited_log2=10)
    dev.search(q, k=10, ef=200, mc=mc, visited="lds", visited_log2=10)
    c = dev.counters()
    assert c["spilled"] > 0 and c["overflowed"] == 0
    print(f"case large ef mc={mc} ef=1000 (Bloom filter + HBM table)", flush=True)
    assert dev.plan(40, k=10, ef=1000, mc=mc)["lds_hash_log2"] == 0
    check_equal(dev, og, q[:40], 10, 1000, mc=mc)


@pytest.mark.parametrize("ef", [400, 512])
def test_large_ef_four_byte_hash_freezes_and_spills(ef):
    """4-byte LDS slots above ef 128 (the plan's choice where 32 KiB costs no
    waves), forced down to 2^11 slots: every query outgrows the hash,
    freezes it and goes on with the HBM table; results, distances and
    counters equal the oracle's, nothing overflows."""
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(100, 128, seed=17)
    plan = dev.plan(100, k=10, ef=ef, visited="lds4", visited_log2=11)
    assert (plan["lds_hash_log2"], plan["lds_hash_2byte"]) == (11, 0), plan
    print(f"case 4-byte LDS hash 2^11 ef={ef} (freeze)", flush=True)
    check_equal(dev, og, q, 10, ef, visited="lds4", visited_log2=11)
    dev.search(q, k=10, ef=ef, visited="lds4", visited_log2=11)
    c = dev.counters()
    assert c["spilled"] > 0 and c["overflowed"] == 0, c


def test_large_ef_full_batch_two_byte_hash():
    """A batch larger than the resident grid at ef=200 takes the 2-byte LDS
    hash (2^13 slots: 4-byte slots in 32 KiB would cost resident waves):
    its results equal the HBM table's for every query and the oracle's on a
    sample."""
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(4000, 128, seed=19)
    plan = dev.plan(4000, k=10, ef=200)
    assert (plan["lds_hash_log2"], plan["lds_hash_2byte"]) == (13, 1), plan
    print("case large ef full batch ef=200", flush=True)
    a = dev.search(q, k=10, ef=200, with_stats=True)
    b = dev.search(q, k=10, ef=200, with_stats=True, visited="hbm")
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
    ids, ds, ost = og.search(q[:300], 10, 200, order=orc.WAVE)
    np.testing.assert_array_equal(a[0][:300], ids.astype(np.int64))
    np.testing.assert_array_equal(a[1][:300], ds)
    np.testing.assert_array_equal(a[2][:300].astype(np.int64), ost)


def test_recall_c1_scaled():
    """100k x 128, M=16, efC=128, k=10, ef=64 (the reference's CPU case):
    GPU recall@10 equals the oracle's (same ids) and is reported."""
    x, g, dev = graph(100000, 128, 16, 128, centres=64, seed=31)
    q = queries(1000, 128, centres=64, seed=31)
    og = orc.OracleGraph(g.arrays())
    lab, dist = dev.search(q, k=10, ef=64)
    ids, ds, _ = og.search(q, 10, 64, order=orc.WAVE)
    np.testing.assert_array_equal(lab, ids)
    gt, _ = orc.bruteforce_f64(x, q, 10)
    r_gpu = orc.recall_at(lab, gt, 10)
    ids_seq, _, _ = og.search(q, 10, 64, order=orc.SEQ)
    r_cpu = orc.recall_at(ids_seq, gt, 10)
    print(f"recall@10 gpu={r_gpu:.4f} cpu(hnswlib order)={r_cpu:.4f}", flush=True)
    assert abs(r_gpu - r_cpu) <= 0.001
    assert r_gpu > 0.9


@pytest.mark.parametrize("metric,k", [(gvs.L2, 10), (gvs.L2, 100), (gvs.IP, 10)])
def test_bruteforce_f64_mfma_matches_oracle(metric, k):
    from gvs import hip as H
    norm = metric == gvs.IP
    x, g, dev = graph(20000, 128, 16, 100, metric=metric, normalize=norm)
    q = queries(150, 128, seed=13, normalize=norm)
    qt = H.DeviceArray.from_host(q)
    ids = H.DeviceArray((150, k), np.int32)
    ds = H.DeviceArray((150, k), np.float64)
    print(f"case bruteforce metric={metric} k={k}", flush=True)
    dev.bruteforce_f64_device(qt.address, 150, 128, k, ids.address, ds.address)
    H.synchronize()
    gi, gd = ids.to_host(), ds.to_host()
    oi, od = orc.bruteforce_f64(x, q, k, metric=metric)
    np.testing.assert_allclose(gd, od, rtol=1e-9, atol=1e-9)
    mism = gi != oi
    for r, c in zip(*np.nonzero(mism)):
        assert abs(gd[r, c] - od[r, c]) <= 1e-9 * max(1.0, abs(od[r, c]))
    assert mism.mean() < 0.01


def test_bruteforce_before_and_after_search_on_one_index():
    """Growing the search scratch must leave the brute-force norms intact."""
    from gvs import hip as H
    x, g, _ = graph(20000, 128, 16, 100)
    dev = gvs.DeviceIndex(g)
    q = queries(120, 128, seed=21)
    oi, od = orc.bruteforce_f64(x, q, 10)
    qt = H.DeviceArray.from_host(q)
    for step in range(2):
        print(f"case bruteforce around search step={step}", flush=True)
        ids = H.DeviceArray((120, 10), np.int32)
        ds = H.DeviceArray((120, 10), np.float64)
        dev.bruteforce_f64_device(qt.address, 120, 128, 10, ids.address, ds.address)
        H.synchronize()
        np.testing.assert_allclose(ds.to_host(), od, rtol=1e-9, atol=1e-9)
        dev.search(queries(3000 * (step + 1), 128, seed=22), k=10, ef=64)


@pytest.mark.parametrize("mc", [1, 4])
def test_lds_visited_spill_is_exact(mc):
    """A query that outgrows a small LDS visited hash freezes it and goes on
    with the HBM table (with and without a Bloom filter in front of it):
    same results and counters as the oracle."""
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(200, 128, seed=15)
    print(f"case lds spill mc={mc}", flush=True)
    check_equal(dev, og, q, 10, 64, mc=mc, visited="lds", visited_log2=10)
    dev.search(q, k=10, ef=64, mc=mc, visited="lds", visited_log2=10)
    c = dev.counters()
    assert c["spilled"] > 0 and c["overflowed"] == 0
    check_equal(dev, og, q, 10, 64, mc=mc, visited="lds", visited_log2=10, bloom_log2=8)
    check_equal(dev, og, q, 10, 64, mc=mc, visited="hbm")
    check_equal(dev, og, q, 10, 64, mc=mc, visited="lds")


def test_async_search_has_no_overflow_at_default_sizes():
    from gvs import hip as H
    x, g, dev = graph(20000, 128, 16, 100)
    q = queries(2000, 128, seed=16)
    dq = H.DeviceArray.from_host(q)
    dl = H.DeviceArray((2000, 10), np.int64)
    dd = H.DeviceArray((2000, 10), np.float32)
    for mc in (1, 2, 4):
        dev.search_device(dq.address, 2000, 128, dl.address, dd.address, k=10, ef=64, mc=mc, asynchronous=True)
        assert dev.overflow_count() == 0
        lab_sync, _ = dev.search(q, k=10, ef=64, mc=mc)
        np.testing.assert_array_equal(dl.to_host(), lab_sync)


def test_fork_batches_on_three_streams_equal_one_handle():
    """gvs_index_fork: handles on the same device arrays with scratch of their
    own.  Batches going round the index, a fork and a fork of the fork on
    three streams (two waves per query, persistent, region order; all queued
    before any is waited on) return the oracle's results, as one handle does
    on one stream; the index cannot be destroyed while a fork is open."""
    from gvs import hip as H
    x, g, _ = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    dev = gvs.DeviceIndex(g)
    fork = dev.fork()
    fork2 = fork.fork()
    handles = (dev, fork, fork2)
    nq = 1000
    q = queries(nq, 128, seed=61)
    cfg = dict(mc=4, waves_per_query=2, persistent=True, region_order=True)
    print("case fork: one handle against the oracle", flush=True)
    lab, dist, _ = check_equal(dev, og, q, 10, 64, **cfg)
    dq = H.DeviceArray.from_host(q)
    streams = [H.Stream() for _ in handles]
    outs = [(H.DeviceArray((nq, 10), np.int64), H.DeviceArray((nq, 10), np.float32)) for _ in range(9)]
    print("case fork: nine batches over three handles and three streams", flush=True)
    for b in range(9):
        handles[b % 3].search_device(dq.address, nq, 128, outs[b][0].address, outs[b][1].address, k=10, ef=64,
                                     stream=streams[b % 3].handle, asynchronous=True, **cfg)
    H.synchronize()
    for b in range(9):
        np.testing.assert_array_equal(outs[b][0].to_host(), lab)
        np.testing.assert_array_equal(outs[b][1].to_host(), dist)
    assert all(h.overflow_count() == 0 for h in handles)
    err_state = -7  # GVS_ERR_STATE
    assert gvs.lib.gvs_index_destroy(dev.handle) == err_state
    assert gvs.lib.gvs_index_destroy(fork.handle) == 0
    fork._h = None
    assert gvs.lib.gvs_index_destroy(dev.handle) == err_state  # fork2 is the index's fork too
    assert gvs.lib.gvs_index_destroy(fork2.handle) == 0
    fork2._h = None
    assert gvs.lib.gvs_index_destroy(dev.handle) == 0
    dev._h = None


@pytest.mark.parametrize("storage", [gvs.BF16, gvs.FP8])
def test_low_precision_traversal_and_fp32_rerank(storage):
    """BF16 / FP8-e4m3 (per-vector scale) storage, fp32 accumulation: the
    traversal equals the oracle on the same quantised rows; the exact fp32
    rerank of the final ef candidates equals the oracle's rerank; distance
    error and recall against the FP64 ground truth are reported."""
    import lowprec
    x = gvs.gen_mixture(20000, 128, 32, 41, 42)
    g = gvs.HostGraph.build_hnsw(x, M=16, ef_construction=100, nthreads=16)
    dev = gvs.DeviceIndex(g, storages=(gvs.F32, gvs.BF16, gvs.FP8))
    a = g.arrays()
    if storage == gvs.BF16:
        og = orc.OracleGraph(a, dim_pad=128, storage=orc.BF16, bf16=lowprec.to_bf16(a["vectors"]))
    else:
        codes, scale = lowprec.quantize_fp8_rows(a["vectors"])
        og = orc.OracleGraph(a, dim_pad=128, storage=orc.FP8, fp8=codes, scale=scale)
    q = gvs.gen_mixture(200, 128, 32, 41, 77)  # same mixture as the base set
    print(f"case low precision storage={storage}", flush=True)
    check_equal(dev, og, q, 10, 64, storage=storage)
    lab, dist = dev.search(q, k=10, ef=64, storage=storage, rerank=True)
    cand, _, _ = og.search(q, 64, 64, order=orc.WAVE)
    ri, rd = orc.rerank_f32(og, q, cand, 10)
    np.testing.assert_array_equal(lab, ri)
    np.testing.assert_array_equal(dist, rd)
    # ef above 128: pools of several 64-entry chunks, the 2-byte LDS visited
    # hash, and a rerank over four chunks of candidates
    print(f"case low precision storage={storage} ef=200", flush=True)
    check_equal(dev, og, q, 10, 200, storage=storage)
    lab2, dist2 = dev.search(q, k=10, ef=200, storage=storage, rerank=True)
    cand2, _, _ = og.search(q, 200, 200, order=orc.WAVE)
    ri2, rd2 = orc.rerank_f32(og, q, cand2, 10)
    np.testing.assert_array_equal(lab2, ri2)
    np.testing.assert_array_equal(dist2, rd2)
    gt, gd = orc.bruteforce_f64(x, q, 10)
    lab32, d32 = dev.search(q, k=10, ef=64)
    r32, rlp = orc.recall_at(lab32, gt, 10), orc.recall_at(lab, gt, 10)
    lab_nr, d_nr = dev.search(q, k=10, ef=64, storage=storage)
    rel = np.abs(d_nr[:, 0] - gd[:, 0]) / gd[:, 0]
    print(f"storage={storage}: recall@10 fp32 {r32:.4f} low-precision+rerank {rlp:.4f} "
          f"(no rerank {orc.recall_at(lab_nr, gt, 10):.4f}); top-1 traversal distance rel. error "
          f"mean {rel.mean():.2e} max {rel.max():.2e}", flush=True)
    assert rlp >= r32 - 0.02


@pytest.mark.parametrize("persistent", [False, True])
def test_captured_graph_replays_match_eager(persistent):
    """A captured search replayed several times, with new queries in its
    buffer and eager searches on the same index between replays (one of
    them large enough to regrow the index's own scratch), gives the eager
    result every time."""
    from gvs import hip as H
    x, g, dev = graph(20000, 128, 16, 100)
    nq, k, ef = 400, 10, 64
    batches = [queries(nq, 128, seed=30 + i) for i in range(3)]
    dq = H.DeviceArray.from_host(batches[0])
    dl = H.DeviceArray((nq, k), np.int64)
    dd = H.DeviceArray((nq, k), np.float32)
    ds = H.DeviceArray((nq, gvs.STATS_ROW), np.int32)
    st = H.Stream()
    sg = dev.capture(dq.address, nq, 128, dl.address, dd.address, k=k, ef=ef, mc=2, stats_ptr=ds.address,
                     persistent=persistent)
    for rep in range(5):
        b = batches[rep % 3]
        print(f"case graph replay persistent={persistent} rep={rep}", flush=True)
        dq.upload(b)
        sg.launch(st.handle)
        st.synchronize()
        lab, dist, cnt = dev.search(b, k=k, ef=ef, mc=2, with_stats=True)
        np.testing.assert_array_equal(dl.to_host(), lab)
        np.testing.assert_array_equal(dd.to_host(), dist)
        np.testing.assert_array_equal(ds.to_host()[:, :gvs.STATS_PER_QUERY], cnt)
        if rep == 2:
            dev.search(queries(6000, 128, seed=40), k=10, ef=200)
    sg.close()


@pytest.mark.parametrize("W,mc,mg", [(2, 1, 1), (2, 2, 1), (4, 4, 1), (8, 8, 1), (8, 2, 1), (4, 8, 1),
                                     (2, 1, 2), (2, 2, 2), (4, 4, 2), (8, 8, 2), (8, 2, 4), (4, 8, 4),
                                     (8, 8, 8), (2, 1, 3)])
def test_intra_query_waves_match_single_wave_dst(W, mc, mg):
    """Several waves per query (one workgroup) give the ids, distances and
    counters of one wave running DST(mc, mg), i.e. the oracle's."""
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(150, 128, seed=50)
    for ef in (32, 128):
        print(f"case intra W={W} mc={mc} mg={mg} ef={ef}", flush=True)
        check_equal(dev, og, q, 10, ef, mc=mc, mg=mg, waves_per_query=W)


@pytest.mark.parametrize("W,mc", [(2, 2), (2, 4), (4, 4), (4, 8)])
def test_intra_query_waves_persistent_and_region_order(W, mc):
    """Several waves per query on the persistent work queue, in arrival order
    and in region order (descent kernel, sort, XCD segments): the ids,
    distances and all four counters of the oracle's DST(mc, 1), with the
    upper-layer counters from the descent kernel."""
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(400, 128, seed=54)
    assert g.arrays()["max_level"] >= 2
    for kw in (dict(persistent=True), dict(persistent=True, region_order=True),
               dict(persistent=True, region_order=True, grid_waves=8 * W)):
        print(f"case intra persistent W={W} mc={mc} {kw}", flush=True)
        check_equal(dev, og, q, 10, 64, mc=mc, waves_per_query=W, **kw)
    check_equal(dev, og, q, 100, 300, mc=mc, waves_per_query=W, persistent=True, region_order=True)
    xi, gi, devi = graph(8000, 96, 12, 100, metric=gvs.IP, normalize=True)
    ogi = orc.OracleGraph(gi.arrays())
    qi = queries(200, 96, seed=55, normalize=True)
    check_equal(devi, ogi, qi, 10, 64, mc=mc, waves_per_query=W, persistent=True, region_order=True)


@pytest.mark.parametrize("W", [2, 8])
def test_intra_query_waves_hbm_spill_ip_and_entries(W):
    """The intra-query path through the HBM visited table (forced at large
    ef), the shared LDS hash above ef 128 (the default) and forced to spill,
    inner product, and given entry points."""
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(100, 128, seed=51)
    print(f"case intra hbm W={W}", flush=True)
    assert dev.plan(100, k=10, ef=200, mc=4, waves_per_query=W)["lds_hash_log2"] == 13
    check_equal(dev, og, q, 10, 200, mc=4, waves_per_query=W)
    check_equal(dev, og, q, 10, 400, mc=4, waves_per_query=W)
    check_equal(dev, og, q, 10, 200, mc=4, waves_per_query=W, visited="hbm")
    check_equal(dev, og, q, 10, 300, mc=4, waves_per_query=W, visited="lds4", visited_log2=10)
    check_equal(dev, og, q, 10, 64, mc=2, waves_per_query=W, visited="lds", visited_log2=10)
    ent = np.random.default_rng(3).integers(0, 20000, size=100).astype(np.int32)
    check_equal(dev, og, q, 10, 64, mc=2, entries=ent, waves_per_query=W)
    xi, gi, devi = graph(8000, 96, 12, 100, metric=gvs.IP, normalize=True)
    ogi = orc.OracleGraph(gi.arrays())
    qi = queries(100, 96, seed=52, normalize=True)
    check_equal(devi, ogi, qi, 10, 64, mc=4, waves_per_query=W)


def test_intra_query_rejects_unsupported():
    x, g, dev = graph(20000, 128, 16, 100)
    q = queries(4, 128, seed=53)
    with pytest.raises(gvs.GvsError):
        dev.search(q, k=10, ef=64, mc=2, mg=9, waves_per_query=4)
    with pytest.raises(gvs.GvsError):
        dev.search(q, k=10, ef=64, mc=2, mg=9)
    with pytest.raises(gvs.GvsError):
        dev.search(q, k=10, ef=64, waves_per_query=3)
    with pytest.raises(gvs.GvsError):
        dev.search(q, k=10, ef=64, waves_per_query=2, storage=gvs.BF16)


@pytest.mark.parametrize("nq", [16, 500])
def test_autotuner_returns_a_working_configuration(nq):
    """gvs_autotune picks a supported configuration; searching with it
    matches the oracle's DST with the same mc and mg."""
    from gvs import hip as H
    x, g, dev = graph(20000, 128, 16, 100)
    og = orc.OracleGraph(g.arrays())
    q = queries(nq, 128, seed=60)
    dq = H.DeviceArray.from_host(q)
    dl = H.DeviceArray((nq, 10), np.int64)
    dd = H.DeviceArray((nq, 10), np.float32)
    print(f"case autotune nq={nq}", flush=True)
    cfg, ms = dev.autotune(dq.address, nq, 128, dl.address, dd.address, k=10, ef=64)
    assert ms > 0 and cfg["ef"] == 64 and cfg["mc"] in (1, 2, 4, 8) and cfg["mg"] in (1, 2, 4)
    assert cfg["waves_per_query"] in ((1, 2, 4, 8) if nq <= 64 else (1, 2, 4))
    kw = {kk: v for kk, v in cfg.items() if kk not in ("ef", "mc", "mg")}
    check_equal(dev, og, q, 10, 64, mc=cfg["mc"], mg=cfg["mg"], **kw)
    # above ef 128 the visited-set placement is one of the timed choices
    print(f"case autotune nq={nq} ef=200", flush=True)
    cfg, ms = dev.autotune(dq.address, nq, 128, dl.address, dd.address, k=10, ef=200)
    assert ms > 0 and cfg["ef"] == 200 and cfg.get("visited") in (None, "hbm")
    kw = {kk: v for kk, v in cfg.items() if kk not in ("ef", "mc", "mg")}
    check_equal(dev, og, q, 10, 200, mc=cfg["mc"], mg=cfg["mg"], **kw)


@pytest.mark.parametrize("mc", [1, 2])
def test_region_order_matches_oracle(mc):
    """Persistent search in region order (descent kernel, sort by region,
    per-XCD queues): ids, distances and all four counters as the oracle."""
    x, g, dev = graph(20000, 128, 16, 100)
    assert g.view.max_level >= 2
    og = orc.OracleGraph(g.arrays())
    q = queries(3000, 128, seed=70)
    print(f"case region order mc={mc}", flush=True)
    check_equal(dev, og, q, 10, 64, mc=mc, persistent=True, region_order=True)
    check_equal(dev, og, q[:700], 10, 200, mc=mc, persistent=True, region_order=True)


@pytest.mark.parametrize("ef", [32, 64, 128, 256])
def test_c2_shape_d96_degree64_matches_oracle(ef):
    """The plain inter-query configuration's shape (D=96, M=32 so layer-0
    degree 64, L2, one batched launch) at reduced N, over its ef sweep."""
    x, g, dev = graph(30000, 96, 32, 200, centres=64)
    og = orc.OracleGraph(g.arrays())
    q = queries(300, 96, centres=64, seed=80)
    print(f"case c2 shape ef={ef}", flush=True)
    check_equal(dev, og, q, 10, ef)
    check_equal(dev, og, q, 10, ef, mc=2, persistent=True, region_order=True)


class ArrayGraph:
    """A gvs_graph view over numpy arrays (flat graph), kept alive here."""

    def __init__(self, a):
        import ctypes as C
        self.a = {kk: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for kk, v in a.items()}
        v = gvs.Graph()
        v.n, v.dim, v.metric = a["n"], a["dim"], a["metric"]
        v.vectors = self.a["vectors"].ctypes.data
        v.ld_vectors = a["dim"]
        v.r0 = a["r0"]
        v.adj0 = self.a["adj0"].ctypes.data
        v.max_level, v.entry, v.ru, v.up_rows = 0, a["entry"], 0, 0
        self.view = v


@pytest.mark.parametrize("mc,mg", [(1, 1), (2, 1), (1, 2), (2, 2)])
def test_dst_hand_traced_fixture_on_gpu(mc, mg):
    """The hand-traced DST fixture (tests/test_oracle.py): the kernel gives
    the derived results, hops and evaluations for each (mc, mg), one wave
    per query and two waves per query."""
    import test_oracle as T
    dev = gvs.DeviceIndex(ArrayGraph(T.dst_fixture_arrays()))
    q = np.zeros((1, 1), dtype=np.float32)
    want_ids, want_groups, want_evals = T.DST_EXPECT[(mc, mg)]
    x = np.float32(T.DST_X)
    runs = [dict(), dict(waves_per_query=2)]
    for kw in runs:
        print(f"case hand-traced DST mc={mc} mg={mg} {kw}", flush=True)
        lab, dist, st = dev.search(q, k=2, ef=2, mc=mc, mg=mg, with_stats=True, **kw)
        assert lab[0].tolist() == want_ids
        np.testing.assert_array_equal(dist[0], x[want_ids] * x[want_ids])
        assert st[0, 0] == sum(len(gr) for gr in want_groups)
        assert st[0, 1] == want_evals


@pytest.mark.parametrize("mc,mg", [(1, 2), (1, 3), (1, 4), (2, 3)])
def test_dst_hand_traced_pipeline_depth_on_gpu(mc, mg):
    """The second hand-traced DST fixture (tests/test_oracle.py), where two
    and three groups in flight give different results: one wave per query
    and two and four waves per query."""
    import test_oracle as T
    dev = gvs.DeviceIndex(ArrayGraph(T.dst2_fixture_arrays()))
    q = np.zeros((1, 1), dtype=np.float32)
    want_ids, want_groups, want_evals = T.DST2_EXPECT[(mc, mg)]
    x = np.float32(T.DST2_X)
    for kw in (dict(), dict(waves_per_query=2), dict(waves_per_query=4)):
        print(f"case hand-traced DST2 mc={mc} mg={mg} {kw}", flush=True)
        lab, dist, st = dev.search(q, k=3, ef=3, mc=mc, mg=mg, with_stats=True, **kw)
        assert lab[0].tolist() == want_ids
        np.testing.assert_array_equal(dist[0], x[want_ids] * x[want_ids])
        assert st[0, 0] == sum(len(gr) for gr in want_groups)
        assert st[0, 1] == want_evals


def test_device_bloom_hash_matches_reference_known_answers():
    """The kernel's MurmurHash2_KeyLen4 (the LDS Bloom filter's hash) equals
    the reference's on the committed known-answer vectors."""
    import os
    from gvs import hip as H
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "murmur2_kat.npz"))
    n = z["keys"].size
    dk = H.DeviceArray.from_host(np.ascontiguousarray(z["keys"], dtype=np.uint32))
    ds = H.DeviceArray.from_host(np.ascontiguousarray(z["seeds"], dtype=np.uint32))
    do = H.DeviceArray((n,), np.uint32)
    print("case device murmur2 known answers", flush=True)
    gvs.murmur2_keylen4_device(dk.address, ds.address, do.address, n)
    H.synchronize()
    np.testing.assert_array_equal(do.to_host(), z["h_keylen4"])


def test_fp8_decode_every_code():
    """Two rows holding every non-NaN OCP e4m3 code once (positive codes in
    row 0, negative in row 1, each row padded with 448 so its scale is 1 and
    every value is its own code): the FP8 traversal's distances to a random
    query equal the oracle's decode (orc_e4m3_to_f32 x scale) bit for bit."""
    import lowprec
    grid = lowprec.GRID  # codes 0..0x7E
    x = np.full((2, 128), 448.0, dtype=np.float32)
    x[0, :127] = grid
    x[1, :127] = -grid
    a = {"n": 2, "dim": 128, "metric": gvs.L2, "r0": 1, "max_level": 0, "entry": 0, "ru": 0, "vectors": x,
         "adj0": np.array([[1], [0]], dtype=np.int32), "upper_off": np.zeros(0, np.int64),
         "adj_up": np.zeros((0, 1), np.int32)}
    codes, scale = lowprec.quantize_fp8_rows(x)
    assert np.all(scale == 1.0) and len(set(codes.ravel().tolist())) == 254
    dev = gvs.DeviceIndex(ArrayGraph(a), storages=(gvs.F32, gvs.FP8))
    og = orc.OracleGraph(a, dim_pad=128, storage=orc.FP8, fp8=codes, scale=scale)
    q = np.random.default_rng(4).normal(0, 30, size=(8, 128)).astype(np.float32)
    print("case fp8 every code", flush=True)
    check_equal(dev, og, q, 2, 2, storage=gvs.FP8)


def ring_index(x, metric=gvs.L2):
    n, dim = x.shape
    a = dict(n=n, dim=dim, metric=metric, vectors=x, r0=1, entry=0,
             adj0=((np.arange(n, dtype=np.int64) + 1) % n).astype(np.int32).reshape(n, 1))
    keep = ArrayGraph(a)
    return keep, gvs.DeviceIndex(keep)


@pytest.mark.parametrize("dim", [20, 64, 96, 128, 160])
@pytest.mark.parametrize("k", [1, 10, 16, 17, 100])
def test_bruteforce_dims_k_ragged_and_ties(dim, k):
    """Every padded width (32/64/96/128 on the register-resident kernel, 160
    on the general one), lists in registers (k <= 16) and in LDS (k > 16), a
    base set and a query count that are not multiples of the tile sizes, and
    exact ties (every vector appears twice): results equal the oracle's FP64
    scan and tied distances come in ascending id order."""
    from gvs import hip as H
    rng = np.random.default_rng(dim * 1000 + k)
    half = rng.standard_normal((2600, dim)).astype(np.float32)
    x = np.concatenate([half, half[:2403]])          # n = 5003, ids i and 2600 + i tie
    q = rng.standard_normal((37, dim)).astype(np.float32)
    metric = gvs.IP if k == 17 else gvs.L2
    keep, dev = ring_index(x, metric)
    qt = H.DeviceArray.from_host(q)
    ids = H.DeviceArray((37, k), np.int32)
    ds = H.DeviceArray((37, k), np.float64)
    print(f"case bruteforce dim={dim} k={k} metric={metric}", flush=True)
    dev.bruteforce_f64_device(qt.address, 37, dim, k, ids.address, ds.address)
    H.synchronize()
    gi, gd = ids.to_host(), ds.to_host()
    oi, od = orc.bruteforce_f64(x, q, k, metric=metric)
    np.testing.assert_allclose(gd, od, rtol=1e-9, atol=1e-9)
    assert np.all(np.diff(gd, axis=1) >= 0)
    tie = np.diff(gd, axis=1) == 0
    assert np.all(np.diff(gi, axis=1)[tie] > 0)
    mism = gi != oi
    for r, c in zip(*np.nonzero(mism)):
        assert abs(gd[r, c] - od[r, c]) <= 1e-9 * max(1.0, abs(od[r, c]))
    assert mism.mean() < 0.02
    if k >= 10:
        assert tie.sum() >= 37 * (k // 4)  # the twins do tie exactly
