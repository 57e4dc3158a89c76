"""CPU tests: the oracle pinned against the reference hnswlib (golden vectors
and oracle/_ref when present), closed forms and invariants; the host-side
builder, formats and generator; the C ABI exports."""
import os
import re
import subprocess

import numpy as np
import pytest

import gvs
from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
CASES = ["l2_n1000_d16_M8", "ip_n1000_d24_M6", "l2_n400_d8_M4"]


def load_case(name):
    z = np.load(os.path.join(GOLD, name + ".npz"))
    g = gvs.HostGraph.load_hnswlib(os.path.join(GOLD, name + ".hnsw.bin"), int(z["x"].shape[1]), int(z["metric"]))
    return z, g


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_golden(name):
    """hnswlib-order oracle == reference searchKnnCloserFirst, bit for bit."""
    z, g = load_case(name)
    og = orc.OracleGraph(g.arrays())
    k = int(z["k"])
    for ef in z["efs"]:
        ids, ds, _ = og.search(z["q"], k, int(ef), order=orc.SEQ)
        np.testing.assert_array_equal(ids, z[f"ids_ef{ef}"])
        np.testing.assert_array_equal(ds, z[f"dists_ef{ef}"])


@pytest.mark.parametrize("name", CASES)
def test_builder_reproduces_reference_graph(name):
    """Single-thread builder in hnswlib's scalar order == the reference graph."""
    z, g = load_case(name)
    ref = g.arrays()
    mine = gvs.HostGraph.build_hnsw(z["x"], M=int(z["M"]), ef_construction=int(z["efc"]), seed=100,
                                    metric=int(z["metric"]), nthreads=1, exact_order=True).arrays()
    assert mine["entry"] == ref["entry"] and mine["max_level"] == ref["max_level"]
    np.testing.assert_array_equal(mine["adj0"], ref["adj0"])
    np.testing.assert_array_equal(mine["upper_off"], ref["upper_off"])
    np.testing.assert_array_equal(mine["adj_up"], ref["adj_up"])
    np.testing.assert_array_equal(mine["vectors"], z["x"])


@pytest.mark.parametrize("name", CASES)
def test_batched_builder_with_batches_of_one_is_the_reference_graph(name):
    """The batched insertion (the GPU builder's rules, run on the host) with
    batches of one node and hnswlib's scalar order == the reference graph."""
    z, g = load_case(name)
    ref = g.arrays()
    mine = gvs.HostGraph.build_hnsw_batched(z["x"], M=int(z["M"]), ef_construction=int(z["efc"]), seed=100,
                                            metric=int(z["metric"]), nthreads=4, exact_order=True,
                                            batch_ratio=-1).arrays()
    assert mine["entry"] == ref["entry"] and mine["max_level"] == ref["max_level"]
    np.testing.assert_array_equal(mine["adj0"], ref["adj0"])
    np.testing.assert_array_equal(mine["upper_off"], ref["upper_off"])
    np.testing.assert_array_equal(mine["adj_up"], ref["adj_up"])


def test_batched_builder_graph_is_sound_and_close_to_sequential():
    """Batches of 5 % of the graph: rows within their width, no self links or
    repeats, every node reachable from the entry, and recall at a fixed ef
    close to the sequential build's."""
    x = gvs.gen_mixture(6000, 32, 24, 61, 62)
    q = gvs.gen_mixture(300, 32, 24, 61, 63)
    gb = gvs.HostGraph.build_hnsw_batched(x, M=12, ef_construction=80, nthreads=8, batch_ratio=0.05)
    gs = gvs.HostGraph.build_hnsw(x, M=12, ef_construction=80, nthreads=1, exact_order=True)
    a = gb.arrays()
    adj = a["adj0"]
    assert adj.shape == (6000, 24)
    for v in range(6000):
        row = adj[v][adj[v] >= 0]
        assert np.all(adj[v][len(row):] == -1) and v not in row and len(set(row.tolist())) == len(row)
    seen, todo = {a["entry"]}, [a["entry"]]
    while todo:
        v = todo.pop()
        for u in adj[v]:
            if u >= 0 and int(u) not in seen:
                seen.add(int(u))
                todo.append(int(u))
    assert len(seen) == 6000
    gt, _ = orc.bruteforce_f64(x, q, 10)
    rb = orc.recall_at(orc.OracleGraph(a).search(q, 10, 32)[0], gt, 10)
    rs = orc.recall_at(orc.OracleGraph(gs.arrays()).search(q, 10, 32)[0], gt, 10)
    assert rb >= rs - 0.02, (rb, rs)


def test_hnswlib_save_roundtrip(tmp_path):
    z, g = load_case(CASES[0])
    p = str(tmp_path / "rt.bin")
    g.save_hnswlib(p)
    g2 = gvs.HostGraph.load_hnswlib(p, int(z["x"].shape[1]))
    a, b = g.arrays(), g2.arrays()
    for key in ("adj0", "upper_off", "adj_up", "vectors", "labels"):
        np.testing.assert_array_equal(a[key], b[key])


@pytest.mark.skipif(not os.path.exists(orc.REF_BIN), reason="oracle/_ref not built (no reference tree)")
def test_oracle_matches_reference_live(tmp_path):
    """Larger seeded case against the reference binary (when it is built)."""
    x = gvs.gen_mixture(4000, 64, 32, 5, 6)
    q = gvs.gen_mixture(200, 64, 32, 5, 7)
    x.tofile(tmp_path / "x.f32")
    q.tofile(tmp_path / "q.f32")
    idx = str(tmp_path / "i.bin")
    subprocess.run([orc.REF_BIN, "build", str(tmp_path / "x.f32"), "4000", "64", "l2", "12", "100", "100", idx],
                   check=True)
    subprocess.run([orc.REF_BIN, "search", idx, "64", "l2", str(tmp_path / "q.f32"), "200", "10", "50",
                    str(tmp_path / "i"), str(tmp_path / "d")], check=True)
    rid = np.fromfile(tmp_path / "i", dtype=np.int32).reshape(200, 10)
    rd = np.fromfile(tmp_path / "d", dtype=np.float32).reshape(200, 10)
    og = orc.OracleGraph(gvs.HostGraph.load_hnswlib(idx, 64).arrays())
    ids, ds, _ = og.search(q, 10, 50, order=orc.SEQ)
    np.testing.assert_array_equal(ids, rid)
    np.testing.assert_array_equal(ds, rd)


def test_wave_order_distance_close_to_f64():
    rng = np.random.default_rng(0)
    for dim in (32, 96, 128, 384):
        a = rng.normal(0, 3, size=(50, dim)).astype(np.float32)
        b = rng.normal(0, 3, size=(50, dim)).astype(np.float32)
        for metric in (orc.L2, orc.IP):
            for i in range(50):
                w = orc.lib.orc_dist_wave(orc._p(a[i]), orc._p(b[i]), dim, metric)
                s = orc.lib.orc_dist_seq(orc._p(b[i]), orc._p(a[i]), dim, metric)
                ref = np.sum((a[i].astype(np.float64) - b[i]) ** 2) if metric == orc.L2 else \
                    1.0 - np.dot(a[i].astype(np.float64), b[i].astype(np.float64))
                scale = max(abs(ref), np.sum(np.abs(a[i].astype(np.float64) * b[i])) if metric else abs(ref))
                assert abs(w - ref) <= 1e-5 * scale + 1e-6
                assert abs(s - ref) <= 1e-5 * scale + 1e-6


@pytest.mark.parametrize("metric", [orc.L2, orc.IP])
def test_fp8_wave_distance_close_to_f64_closed_form(metric):
    """FP8 rows (code * per-vector scale): the kernels'-order distance is
    within 1e-5 relative of sum((e4m3(c) * s - q)^2) (L2) or 1 - sum(e4m3(c)
    * s * q) (IP) in FP64; ignoring the scale or a code fails by far."""
    import lowprec
    rng = np.random.default_rng(12)
    x = rng.normal(0, 3, size=(64, 256)).astype(np.float32)
    codes, scale = lowprec.quantize_fp8_rows(x)
    a = {"n": 64, "dim": 256, "metric": metric, "r0": 1, "max_level": 0, "entry": 0, "ru": 0, "vectors": x,
         "adj0": np.zeros((64, 1), np.int32), "upper_off": np.zeros(0, np.int64), "adj_up": np.zeros((0, 1), np.int32)}
    og = orc.OracleGraph(a, dim_pad=256, storage=orc.FP8, fp8=codes, scale=scale)
    dec = np.array([orc.lib.orc_e4m3_to_f32(int(c)) for c in range(256)], dtype=np.float64)
    for j in range(8):
        q = rng.normal(0, 3, size=256).astype(np.float32)
        for i in range(64):
            xv = dec[codes[i]] * np.float64(scale[i])
            ref = np.sum((xv - q) ** 2) if metric == orc.L2 else 1.0 - np.dot(xv, q.astype(np.float64))
            mag = abs(ref) if metric == orc.L2 else np.sum(np.abs(xv * q))
            got = orc.lib.orc_dist(C_byref(og.g), orc._p(q), i, orc.WAVE)
            assert abs(got - ref) <= 1e-5 * max(mag, 1e-6), (i, got, ref)


def C_byref(s):
    import ctypes
    return ctypes.byref(s)


def test_wave_order_is_the_stated_butterfly():
    """Closed form of the summation order on a vector where order matters."""
    dim = 32
    a = np.zeros(dim, dtype=np.float32)
    # chunk c = lanes p = c % 8; give lane partials very different magnitudes
    for c in range(8):
        a[4 * c] = np.float32(2.0 ** (12 - 3 * c))
    b = np.zeros(dim, dtype=np.float32)
    s = [np.float32(a[4 * p]) ** 2 for p in range(8)]
    f = np.float32
    expect = (f(s[0] + s[4]) + f(s[2] + s[6])) + (f(s[1] + s[5]) + f(s[3] + s[7]))
    got = orc.lib.orc_dist_wave(orc._p(a), orc._p(b), dim, orc.L2)
    assert np.float32(got) == np.float32(expect)


@pytest.mark.parametrize("name", CASES)
def test_wave_order_search_matches_seq_except_ties(name):
    z, g = load_case(name)
    og = orc.OracleGraph(g.arrays())
    k = int(z["k"])
    ef = int(z["efs"][1])
    i1, d1, _ = og.search(z["q"], k, ef, order=orc.SEQ)
    i2, d2, _ = og.search(z["q"], k, ef, order=orc.WAVE)
    np.testing.assert_allclose(d2, d1, rtol=1e-5, atol=1e-6)
    diff = i1 != i2
    # any id difference must sit on a distance tie (within 1e-5 relative)
    for r, c in zip(*np.nonzero(diff)):
        assert abs(d1[r, c] - d2[r, c]) <= 1e-5 * max(abs(d1[r, c]), 1e-6)


def test_exhaustive_ef_gives_exact_knn():
    z, g = load_case("l2_n400_d8_M4")
    og = orc.OracleGraph(g.arrays())
    ids, ds, _ = og.search(z["q"], 5, 400, order=orc.SEQ)
    bi, bd = orc.bruteforce_f64(z["x"], z["q"], 5)
    np.testing.assert_array_equal(ids, bi)
    np.testing.assert_allclose(ds, bd, rtol=1e-5)


def test_bruteforce_matches_numpy():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(700, 20)).astype(np.float32)
    q = rng.normal(size=(30, 20)).astype(np.float32)
    for metric in (orc.L2, orc.IP):
        ids, ds = orc.bruteforce_f64(x, q, 7, metric=metric)
        x64, q64 = x.astype(np.float64), q.astype(np.float64)
        if metric == orc.L2:
            d = ((q64[:, None, :] - x64[None]) ** 2).sum(-1)
        else:
            d = 1.0 - q64 @ x64.T
        order = np.lexsort((np.broadcast_to(np.arange(700), d.shape), d), axis=1)[:, :7]
        np.testing.assert_array_equal(ids, order)
        np.testing.assert_allclose(ds, np.take_along_axis(d, order, 1), rtol=1e-12, atol=1e-12)


def test_dst_mc1_mg1_equals_plain_and_invariants():
    x = gvs.gen_mixture(6000, 32, 32, 21, 22)
    q = gvs.gen_mixture(300, 32, 32, 21, 23)
    g = gvs.HostGraph.build_hnsw(x, M=12, ef_construction=80, nthreads=8)
    og = orc.OracleGraph(g.arrays())
    ids0, d0, st0 = og.search(q, 10, 48)
    ids1, d1, st1 = og.search(q, 10, 48, mc=1, mg=1)
    np.testing.assert_array_equal(ids0, ids1)
    np.testing.assert_array_equal(st0, st1)
    gt, _ = orc.bruteforce_f64(x, q, 10)
    r0 = orc.recall_at(ids0, gt, 10)
    for mc, mg in ((2, 1), (4, 1), (1, 2), (2, 2), (8, 1)):
        i, d, st = og.search(q, 10, 48, mc=mc, mg=mg)
        assert orc.recall_at(i, gt, 10) >= r0 - 0.01
        # delayed synchronisation evaluates at least as many nodes on average
        assert st[:, 1].mean() >= st0[:, 1].mean() * 0.99
        assert np.all(np.diff(d, axis=1) >= 0)


def test_murmur2_matches_reference_known_answers():
    """orc_murmur2 / orc_murmur2_keylen4 equal the reference's MurmurHash2 and
    MurmurHash2_KeyLen4 (unit_tests/bloom_and_hash/sw_murmur2/murmur2.cpp:8-90,
    compiled from the reference tree into oracle/_ref/ref_murmur2, which
    wrote tests/golden/murmur2_kat.npz) on 1,000 (key, seed) pairs and on
    byte strings of every tail length 0..11."""
    z = np.load(os.path.join(GOLD, "murmur2_kat.npz"))
    for key, seed, hg, h4 in zip(z["keys"], z["seeds"], z["h_generic4"], z["h_keylen4"]):
        b = np.array([key], dtype=np.uint32)
        assert orc.lib.orc_murmur2(orc._p(b), 4, int(seed)) == int(hg)
        assert orc.lib.orc_murmur2_keylen4(int(key), int(seed)) == int(h4)
    assert np.array_equal(z["h_generic4"], z["h_keylen4"])  # the reference's own invariant
    for n, blob, seed, h in zip(z["lens"], z["blobs"], z["blob_seeds"], z["h_blobs"]):
        b = np.ascontiguousarray(blob[:max(int(n), 1)])
        assert orc.lib.orc_murmur2(orc._p(b), int(n), int(seed)) == int(h)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(orc.REF_BIN), "ref_murmur2")),
                    reason="oracle/_ref not built (no reference tree)")
def test_murmur2_known_answers_match_live_reference():
    """The committed vectors are what the reference source computes now."""
    z = np.load(os.path.join(GOLD, "murmur2_kat.npz"))
    lines = [f"4 {int(s)} {np.array([k], dtype='<u4').tobytes().hex()}" for k, s in zip(z["keys"][:50], z["seeds"][:50])]
    out = subprocess.run([os.path.join(os.path.dirname(orc.REF_BIN), "ref_murmur2")], input="\n".join(lines) + "\n",
                         capture_output=True, text=True, check=True).stdout.split()
    assert [int(v) for v in out[0::2]] == z["h_generic4"][:50].tolist()


# A hand-traced DST fixture.  Eight nodes on a line, query at x = 0, L2
# (d = x^2), flat graph, entry E, ef = k = 2:
#   id: 0 E x=10 (d=100)  1 A x=6 (36)  2 B x=7 (49)  3 C x=9 (81)
#       4 D x=5.5 (30.25)  5 F x=1 (1)  6 H x=0.5 (0.25)  7 I x=5.2 (27.04)
#   links: E->A,B,C  A->E,D  B->E,F  C->E  D->A,I  F->B,H  H->F  I->D
# F (and H behind it) hide behind B, which plain search never expands.
# Semantics: a group pops up to mc candidates from the candidate set while
# the closest is not beyond the worst of the ef results, evaluates their
# unvisited neighbours at once; up to mg groups are in flight, and the
# oldest completes (its results are merged) before the pipeline is refilled
# (vector_search_baselines/unused/hnswalg.h:343-485, INTER/scheduler.hpp:100-157).
DST_X = [10.0, 6.0, 7.0, 9.0, 5.5, 1.0, 0.5, 5.2]
DST_ADJ = [[1, 2, 3], [0, 4], [0, 5], [0], [1, 7], [2, 6], [5], [4]]
E, A, B, C, D, F, H, I = range(8)
DST_EXPECT = {
    # plain: expand E (A, B, C evaluated; top {A 36, B 49}), A (D; top
    # {D, A}), D (30.25 <= 36; I; top {I, D}), I (D visited); B = 49 is
    # beyond the worst result 30.25: stop.  6 evaluations, 4 hops.
    (1, 1): ([I, D], [[E], [A], [D], [I]], 6),
    # mc = 2: E; then A and B together (B = 49 is not beyond the worst, 49),
    # D and F merged (top {F 1, D 30.25}); then F and D together (H, I; top
    # {H, F}); then H.  8 evaluations, 6 hops, 4 groups.
    (2, 1): ([H, F], [[E], [A, B], [F, D], [H]], 8),
    # mg = 2: E; A is popped, and B is popped before A's group merges (the
    # delayed threshold is still 49); A's group merges (top {D, A}); D is
    # popped before B's group merges, then B's (top {F, D}); F popped, D's
    # group merges I (top {F, I}); I popped, F's group merges H (top
    # {H, F}); H popped.  8 evaluations, 7 hops, 7 groups.
    (1, 2): ([H, F], [[E], [A], [B], [D], [F], [I], [H]], 8),
    # mc = 2, mg = 2: after each merge the candidate set holds at most two
    # entries, so the second group in flight is always empty and the trace
    # is that of (2, 1).
    (2, 2): ([H, F], [[E], [A, B], [F, D], [H]], 8),
}


def dst_fixture_arrays():
    adj = -np.ones((8, 4), dtype=np.int32)
    for v, row in enumerate(DST_ADJ):
        adj[v, :len(row)] = row
    x = np.zeros((8, 1), dtype=np.float32)
    x[:, 0] = DST_X
    return {"n": 8, "dim": 1, "metric": orc.L2, "r0": 4, "max_level": 0, "entry": E, "ru": 0, "vectors": x,
            "adj0": adj, "upper_off": np.zeros(0, np.int64), "adj_up": np.zeros((0, 1), np.int32)}


@pytest.mark.parametrize("mc,mg", sorted(DST_EXPECT))
def test_dst_hand_traced_expansion_order(mc, mg):
    og = orc.OracleGraph(dst_fixture_arrays())
    q = np.zeros(1, dtype=np.float32)
    ids, ds, st, groups = og.trace(q, 2, 2, mc=mc, mg=mg)
    want_ids, want_groups, want_evals = DST_EXPECT[(mc, mg)]
    assert ids.tolist() == want_ids
    x = np.float32(DST_X)
    np.testing.assert_array_equal(ds, x[want_ids] * x[want_ids])
    assert groups == want_groups
    assert st[0] == sum(len(gr) for gr in want_groups)        # hops
    assert st[1] == want_evals                                 # evaluations (entry included)
    assert st[4] == len(want_groups)                           # groups


# A second hand-traced fixture, where the depth of the pipeline matters.
# Ten nodes on a line, query at x = 0, L2, flat graph, entry P0, ef = k = 3:
#   id:  0 x=10 (d=100)  1 x=4.5 (20.25)  2 x=9.5 (90.25)  3 x=5.5 (30.25)
#        4 x=8 (64)      5 x=2.5 (6.25)   6 x=7.5 (56.25)  7 x=8.5 (72.25)
#        8 x=7 (49)      9 x=5 (25)
#   links: 0->2,6,4  1->3,9  2->5,0,3  3->6,0  4->8,9  5->7  6->1,0,9
#          7->6,9  8->6,9  9->4,6
# Node 5 (the nearest) hides behind node 2, the worst of the first three
# results; only a threshold that lags two groups behind lets 2 be expanded.
DST2_X = [10.0, 4.5, 9.5, 5.5, 8.0, 2.5, 7.5, 8.5, 7.0, 5.0]
DST2_ADJ = [[2, 6, 4], [3, 9], [5, 0, 3], [6, 0], [8, 9], [7], [1, 0, 9], [6, 9], [6, 9], [4, 6]]
DST2_EXPECT = {
    # Every run starts: group {0} evaluates 2, 6, 4 (4 evaluations); merged,
    # the results are {6 56.25, 4 64, 2 90.25}.
    # mg = 2: groups {6} (evaluates 1, 9) and {4} (8) go out.  {6} merges
    # (results {1, 9, 6}, worst 56.25); {1} goes out (3).  {4} merges ({1, 9,
    # 8}, worst 49); {9} goes out (nothing new).  {1} merges ({1, 9, 3},
    # worst 30.25); {3} goes out (nothing).  Node 2 (90.25) and node 8 (49)
    # are beyond 30.25: stop.  8 evaluations.
    (1, 2): ([1, 9, 3], [[0], [6], [4], [1], [9], [3]], 8),
    # mg = 3: {6}, {4} and {2} go out while the worst result is still 90.25;
    # {2} evaluates 5 and 3.  {6} merges ({6, 1, 9}, worst 56.25); {1} goes
    # out (3 and 9 already visited).  {4} merges ({1, 9, 8}, worst 49); {9}
    # goes out (nothing).  {2} merges 5 ({5, 1, 9}, worst 25; 3 at 30.25
    # stays out); {5} goes out (7 at 72.25).  8 is beyond 25: stop.
    # 10 evaluations.
    (1, 3): ([5, 1, 9], [[0], [6], [4], [2], [1], [9], [5]], 10),
    # mg = 4: {6}, {4}, {2} go out (the candidate set is then empty).  {6}
    # merges ({6, 1, 9}); {1} and {9} go out (nothing new).  {4} merges
    # ({1, 9, 8}, worst 49); {8} goes out (nothing).  {2} merges ({5, 1, 9});
    # {5} goes out (7).  10 evaluations, 8 groups.
    (1, 4): ([5, 1, 9], [[0], [6], [4], [2], [1], [9], [8], [5]], 10),
    # mc = 2, mg = 3: {6, 4} (evaluates 1, 9, 8) then {2} (5, 3).  {6, 4}
    # merges ({1, 9, 8}, worst 49); {1, 9} and {8} go out (nothing new).  {2}
    # merges ({5, 1, 9}, worst 25); {5} goes out (7).  10 evaluations.
    (2, 3): ([5, 1, 9], [[0], [6, 4], [2], [1, 9], [8], [5]], 10),
}


def dst2_fixture_arrays():
    adj = -np.ones((10, 3), dtype=np.int32)
    for v, row in enumerate(DST2_ADJ):
        adj[v, :len(row)] = row
    x = np.zeros((10, 1), dtype=np.float32)
    x[:, 0] = DST2_X
    return {"n": 10, "dim": 1, "metric": orc.L2, "r0": 3, "max_level": 0, "entry": 0, "ru": 0, "vectors": x,
            "adj0": adj, "upper_off": np.zeros(0, np.int64), "adj_up": np.zeros((0, 1), np.int32)}


@pytest.mark.parametrize("mc,mg", sorted(DST2_EXPECT))
def test_dst_hand_traced_pipeline_depth(mc, mg):
    og = orc.OracleGraph(dst2_fixture_arrays())
    q = np.zeros(1, dtype=np.float32)
    ids, ds, st, groups = og.trace(q, 3, 3, mc=mc, mg=mg)
    want_ids, want_groups, want_evals = DST2_EXPECT[(mc, mg)]
    assert ids.tolist() == want_ids
    x = np.float32(DST2_X)
    np.testing.assert_array_equal(ds, x[want_ids] * x[want_ids])
    assert groups == want_groups
    assert st[0] == sum(len(gr) for gr in want_groups)
    assert st[1] == want_evals
    assert st[4] == len(want_groups)


def test_dst_group_count_below_hops_for_multi_candidate_groups():
    x = gvs.gen_mixture(6000, 32, 32, 21, 22)
    q = gvs.gen_mixture(300, 32, 32, 21, 23)
    g = gvs.HostGraph.build_hnsw(x, M=12, ef_construction=80, nthreads=8)
    og = orc.OracleGraph(g.arrays())
    _, _, s1 = og.search(q, 10, 48, with_groups=True)
    assert np.array_equal(s1[:, 4], s1[:, 0])                 # mc = 1: one hop per group
    for mc in (2, 4, 8):
        _, _, s = og.search(q, 10, 48, mc=mc, with_groups=True)
        assert np.all(s[:, 4] <= s[:, 0]) and np.all(s[:, 0] <= mc * s[:, 4])
        assert s[:, 4].mean() < 0.8 * s[:, 0].mean()


def test_rerank_matches_numpy_sort():
    """orc_rerank_f32: distances of the candidates (orc_dist_wave, pinned
    above) sorted by (distance, id) with numpy, first k."""
    rng = np.random.default_rng(8)
    base = rng.normal(0, 2, size=(500, 64)).astype(np.float32)
    base[7] = base[3]                                           # a distance tie, broken by id
    q = (base[3] + rng.normal(0, 0.05, size=64)).astype(np.float32)  # 3 and 7 tie at the top
    cand = np.concatenate([[7, 3], rng.choice(500, 60, replace=False)]).astype(np.int32)
    cand = np.unique(cand)[rng.permutation(np.unique(cand).size)].astype(np.int32)
    for metric in (orc.L2, orc.IP):
        d = np.array([orc.lib.orc_dist_wave(orc._p(base[c]), orc._p(q), 64, metric) for c in cand], np.float32)
        order = np.lexsort((cand, d))[:12]
        ids = np.empty(12, np.int32)
        ds = np.empty(12, np.float32)
        n = orc.lib.orc_rerank_f32(orc._p(base), 64, metric, orc._p(q), orc._p(cand), cand.size, 12, orc.WAVE,
                                   orc._p(ids), orc._p(ds))
        assert n == 12
        np.testing.assert_array_equal(ids, cand[order])
        np.testing.assert_array_equal(ds, d[order])
        if metric == orc.L2:
            assert ids[:2].tolist() == [3, 7]


def test_bloom_no_false_negatives_and_fp_rate():
    nwords = 1 << 10
    words = np.zeros(nwords, dtype=np.uint64)
    ins = 3000
    for i in range(ins):
        assert orc.lib.orc_bloom_test_and_set(orc._p(words), nwords, 7, i * 7919) == 0 or True
    for i in range(ins):
        assert orc.lib.orc_bloom_test_and_set(orc._p(words), nwords, 7, i * 7919) == 1
    fp = 0
    trials = 100000
    w2 = words.copy()
    for i in range(trials):
        fp += orc.lib.orc_bloom_test_and_set(orc._p(w2), nwords, 7, 10**8 + i)
        w2[:] = words
    rate = fp / trials
    # blocked filter estimate: per-word Poisson load lambda, 6 bits of 64
    lam = ins / nwords
    j = np.arange(0, 40)
    from math import factorial, exp
    pois = np.array([exp(-lam) * lam ** int(t) / factorial(int(t)) for t in j])
    est = float(np.sum(pois * (1 - (1 - 1 / 64) ** (6 * j)) ** 6))
    assert 0.5 * est <= rate <= 1.5 * est + 1e-3


def test_low_precision_encoders_match_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(0, 50, 4000), rng.normal(0, 0.01, 2000), [0.0, -0.0, 447.9, -448.0]]).astype(np.float32)
    tb = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().astype(np.uint16)
    ob = np.array([orc.lib.orc_f32_to_bf16(float(v)) for v in x], dtype=np.uint16)
    np.testing.assert_array_equal(ob, tb)
    xs = np.clip(x, -448, 448)
    t8 = torch.from_numpy(xs).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    o8 = np.array([orc.lib.orc_f32_to_e4m3(float(v)) for v in xs], dtype=np.uint8)
    np.testing.assert_array_equal(o8, t8)
    back = torch.from_numpy(t8).view(torch.float8_e4m3fn).to(torch.float32).numpy()
    ob8 = np.array([orc.lib.orc_e4m3_to_f32(int(v)) for v in t8], dtype=np.float32)
    np.testing.assert_array_equal(ob8, back)


def test_gen_mixture_deterministic_and_shaped():
    a = gvs.gen_mixture(5000, 16, 8, 1, 2, nthreads=1)
    b = gvs.gen_mixture(5000, 16, 8, 1, 2, nthreads=7)
    np.testing.assert_array_equal(a, b)
    c = gvs.gen_mixture(5000, 16, 8, 1, 3)
    assert not np.array_equal(a, c)
    # per-point noise ~ N(0,1): within-cluster spread, centres ~ N(0, 16)
    assert 14.0 < np.var(a, axis=0).mean() < 20.0
    u = gvs.gen_mixture(100, 16, 8, 1, 2, normalize=True)
    np.testing.assert_allclose(np.linalg.norm(u, axis=1), 1.0, rtol=1e-5)


def test_cabi_exports_every_declared_symbol():
    hdr = open(os.path.join(os.path.dirname(HERE), "include", "gvs.h")).read()
    declared = set(re.findall(r"GVS_API\s+[\w\s\*]+?\b(gvs_\w+)\s*\(", hdr))
    assert len(declared) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", gvs.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT\s+(gvs_\w+)", out))
    missing = declared - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"


def test_nsg_format_reader(tmp_path):
    x = gvs.gen_mixture(300, 8, 4, 1, 2)
    rng = np.random.default_rng(4)
    width, ep = 6, 17
    rows = [rng.choice(300, size=rng.integers(1, width + 1), replace=False) for _ in range(300)]
    with open(tmp_path / "g.nsg", "wb") as f:
        np.array([width, ep], dtype=np.uint32).tofile(f)
        for r in rows:
            np.array([len(r)], dtype=np.uint32).tofile(f)
            r.astype(np.uint32).tofile(f)
    g = gvs.HostGraph.load_nsg(str(tmp_path / "g.nsg"), x)
    a = g.arrays()
    assert a["entry"] == ep and a["max_level"] == 0 and a["r0"] == width
    for i, r in enumerate(rows):
        np.testing.assert_array_equal(a["adj0"][i, : len(r)], r)
        assert np.all(a["adj0"][i, len(r):] == -1)


def test_invalid_arguments_return_errors():
    with pytest.raises(gvs.GvsError):
        gvs.HostGraph.load_hnswlib("/nonexistent/file.bin", 16)
    with pytest.raises(gvs.GvsError):
        gvs.HostGraph.build_hnsw(np.zeros((10, 4), np.float32), M=1)
    # index handles: NULL is refused before any device call
    import ctypes as C
    h = C.c_void_p()
    assert gvs.lib.gvs_index_fork(None, C.byref(h)) == -1 and not h.value  # GVS_ERR_ARG
    assert gvs.lib.gvs_index_destroy(None) == -1


def test_vectorised_test_encoders_match_oracle():
    import lowprec
    rng = np.random.default_rng(9)
    x = np.concatenate([rng.normal(0, 40, 20000), rng.normal(0, 0.003, 5000),
                        [0.0, -0.0, 447.0, -447.9, 448.0, 1e-7]]).astype(np.float32)
    ob = np.array([orc.lib.orc_f32_to_bf16(float(v)) for v in x], dtype=np.uint16)
    np.testing.assert_array_equal(lowprec.to_bf16(x), ob)
    xs = np.clip(x, -448, 448)
    o8 = np.array([orc.lib.orc_f32_to_e4m3(float(v)) for v in xs], dtype=np.uint8)
    np.testing.assert_array_equal(lowprec.to_e4m3(xs), o8)


def test_nsg_style_flat_graph_keeps_long_range_edges():
    """to_nsg(R): rows of <= R ids, every node of level >= 1 keeps (up to
    R/2 of) its upper-layer links, the entry is such a node, and greedy
    search from that single entry finds the true neighbours."""
    x = gvs.gen_mixture(20000, 32, 64, 7, 8)
    q = gvs.gen_mixture(200, 32, 64, 7, 9)
    g = gvs.HostGraph.build_hnsw(x, M=8, ef_construction=100, nthreads=8)
    a = g.arrays()
    f = g.to_nsg(16).arrays()
    assert f["max_level"] == 0 and f["r0"] == 16
    lev = np.zeros(a["n"], dtype=np.int64)
    up = a["upper_off"]
    has_up = up >= 0
    # levels from consecutive upper rows
    offs = np.sort(up[has_up])
    ends = np.append(offs[1:], a["adj_up"].shape[0])
    lev_by_off = dict(zip(offs.tolist(), (ends - offs).tolist()))
    for v in np.nonzero(has_up)[0]:
        lev[v] = lev_by_off[int(up[v])]
    assert lev[f["entry"]] >= 1
    for v in np.nonzero(lev >= 1)[0][:300]:
        row = set(f["adj0"][v][f["adj0"][v] >= 0].tolist())
        links = []
        for l in range(int(lev[v]), 0, -1):
            r = a["adj_up"][up[v] + l - 1]
            links += [u for u in r.tolist() if u >= 0 and u != v and u not in links]
        assert set(links[:8]) <= row
    og = orc.OracleGraph(f)
    ids, _, _ = og.search(q, 10, 64)
    gt, _ = orc.bruteforce_f64(x, q, 10)
    assert orc.recall_at(ids, gt, 10) > 0.95


@pytest.mark.parametrize("nc", [1, 4])
def test_fpga_layout_roundtrip_and_bytes(tmp_path, nc):
    """gvs_fpga_save / gvs_fpga_load: the reference's FPGA-layout directory
    (scripts_hnsw/hnsw.py:264-377, scripts_nsg/nsg_to_FPGA.py:56-137):
    byte layout of the records, and a lossless round trip (HNSW and NSG)."""
    x = gvs.gen_mixture(3000, 24, 16, 5, 6)
    g = gvs.HostGraph.build_hnsw(x, M=6, ef_construction=60, nthreads=4)
    a = g.arrays()
    d = tmp_path / "hnsw"
    d.mkdir()
    g.save_fpga(d, nc)
    meta = np.fromfile(d / "meta.bin", dtype=np.uint32)
    assert meta.tolist() == [a["n"], a["max_level"], a["entry"], a["ru"], a["r0"]]
    lrec = (a["r0"] + 15) // 16 * 64 + 64
    vrec = (24 + 15) // 16 * 64 + 64
    l0 = np.fromfile(d / f"ground_links_{nc}_chan_0.bin", dtype=np.uint8).reshape(-1, lrec)
    v0 = np.fromfile(d / f"ground_vectors_{nc}_chan_0.bin", dtype=np.uint8).reshape(-1, vrec)
    assert l0.shape[0] == (a["n"] + nc - 1) // nc
    for j in (0, 1, 7):
        i = j * nc  # channel 0 holds nodes 0, nc, 2nc, ...
        row = a["adj0"][i]
        cnt = int((row >= 0).sum())
        rec = l0[j]
        assert rec[:4].view(np.uint32)[0] == cnt and not rec[4:64].any()
        assert (rec[64:64 + 4 * cnt].view(np.int32) == row[:cnt]).all()
        assert (rec[64:64 + 4 * a["r0"]].view(np.int32)[cnt:] == 0).all()
        vr = v0[j]
        assert (vr[:4 * 24].view(np.float32) == x[i]).all()
        assert vr[(24 + 15) // 16 * 64:][:4].view(np.int32)[0] == -1
    h = gvs.HostGraph.load_fpga(d, 0, 24, nc).arrays()
    for key in ("n", "r0", "max_level", "entry", "ru"):
        assert h[key] == a[key], key
    for key in ("adj0", "upper_off", "adj_up", "vectors"):
        np.testing.assert_array_equal(h[key], a[key])
    q = gvs.gen_mixture(50, 24, 16, 5, 7)
    i1 = orc.OracleGraph(a).search(q, 5, 32)[0]
    i2 = orc.OracleGraph(h).search(q, 5, 32)[0]
    np.testing.assert_array_equal(i1, i2)
    # NSG-style flat graph: kind 1, meta = n, entry, width
    f = g.to_nsg(8)
    fd = tmp_path / "nsg"
    fd.mkdir()
    f.save_fpga(fd, nc)
    fa = f.arrays()
    assert np.fromfile(fd / "meta.bin", dtype=np.uint32).tolist() == [fa["n"], fa["entry"], 8]
    fh = gvs.HostGraph.load_fpga(fd, 1, 24, nc).arrays()
    np.testing.assert_array_equal(fh["adj0"], fa["adj0"])
    assert fh["entry"] == fa["entry"] and fh["max_level"] == 0
    with pytest.raises(gvs.GvsError):
        gvs.HostGraph.load_fpga(fd, 1, 24, nc + 1)  # wrong channel count: files missing


def test_net_frame_sizes_follow_the_reference_packets():
    """Wire sizes of the network front end: 64-byte packets of 16 values,
    as networked_FPGA/CPU_programs/CPU_client.cpp:183-203 sizes them."""
    from gvs import net
    for dim, ef in ((128, 64), (100, 10), (96, 16), (384, 512), (1, 1)):
        hdr, qb, rb = net.frame_bytes(dim, ef)
        assert hdr == 64
        assert qb == 64 * (dim // 16 + (dim % 16 != 0))
        assert rb == 64 + 2 * 64 * (ef // 16 + (ef % 16 != 0))


# NSG search from an initial pool (nsg.py:62-125), hand-traced on the second
# DST fixture with the initial pool [7, 8, 4] (as if the navigating node's
# first three neighbours), L = K = 3:
#   the pool is evaluated (3 evaluations): retset [8 49, 4 64, 7 72.25];
#   expand 8: 6 (56.25) enters at 1, 9 (25) at 0 -> [9, 8, 6] (+ 4 in the
#     L+1-th place); k goes back to 0;
#   expand 9: 4 and 6 already flagged; k moves on past 8 (expanded) to 2;
#   expand 6: 1 (20.25) enters at 0 -> [1, 9, 8]; 0 (100) is not below 49;
#     9 flagged; k back to 0;
#   expand 1: 3 (30.25) enters at 2 -> [1, 9, 3]; k moves to 1 (9 expanded),
#     then 2;
#   expand 3: 6 and 0 flagged; k reaches L: stop.
# Results [1, 9, 3], 5 expansions, 3 + 5 = 8 evaluations.  Node 5 (the
# nearest) sits behind node 2, which is never evaluated.
NSG_INIT = [7, 8, 4]
NSG_EXPECT = ([1, 9, 3], 5, 8)


def test_nsg_search_hand_traced():
    og = orc.OracleGraph(dst2_fixture_arrays())
    q = np.zeros((1, 1), dtype=np.float32)
    want_ids, want_hops, want_evals = NSG_EXPECT
    x = np.float32(DST2_X)
    for fn in (lambda: orc.nsg_search(og, q, 3, 3, NSG_INIT), lambda: orc.search_pool(og, q, 3, 3, NSG_INIT)):
        ids, ds, st = fn()
        assert ids[0].tolist() == want_ids
        np.testing.assert_array_equal(ds[0], x[want_ids] * x[want_ids])
        assert st[0].tolist() == [want_hops, want_evals]


def test_nsg_search_equals_pool_search_and_seeded_hnswlib_search():
    """Two independent restatements agree: nsg.py's sorted-retset walk and
    the two-heap layer-0 search seeded with the same pool (ids, distances,
    expansions, evaluations), over random pools with repeats and ids out of
    range; with the pool {ep} both equal the hnswlib search from ep, which
    is pinned to the reference hnswlib."""
    x = gvs.gen_mixture(3000, 24, 16, 31, 32)
    q = gvs.gen_mixture(80, 24, 16, 31, 33)
    g = gvs.HostGraph.build_hnsw(x, M=8, ef_construction=60, nthreads=8)
    nsg = g.to_nsg(24)
    og = orc.OracleGraph(nsg.arrays())
    rng = np.random.default_rng(5)
    for L in (10, 32, 70):
        init = rng.integers(-5, 3005, size=(80, L)).astype(np.int32)
        init[:, 3] = init[:, 1]  # a repeat
        a = orc.nsg_search(og, q, 10, L, init)
        b = orc.search_pool(og, q, 10, L, init)
        for u, v in zip(a, b):
            np.testing.assert_array_equal(u, v)
    ep = np.full((80, 1), nsg.view.entry, dtype=np.int32)
    for L in (16, 48):
        ids, ds, st = orc.nsg_search(og, q, 10, L, ep)
        ids2, ds2, st2 = og.search(q, 10, L, entries=ep[:, 0])
        np.testing.assert_array_equal(ids, ids2)
        np.testing.assert_array_equal(ds, ds2)
        np.testing.assert_array_equal(st[:, 0], st2[:, 0])
        np.testing.assert_array_equal(st[:, 1], st2[:, 1])


def test_nsg_search_exhaustive_pool_gives_exact_knn():
    """L >= N evaluates every node reachable from the pool: the result is
    the exact kNN among those nodes (FP64 brute force over the set a
    breadth-first search reaches)."""
    x = gvs.gen_mixture(400, 8, 4, 41, 42)
    q = gvs.gen_mixture(20, 8, 4, 41, 43)
    g = gvs.HostGraph.build_hnsw(x, M=6, ef_construction=40, nthreads=1, exact_order=True)
    a = g.to_nsg(12).arrays()
    og = orc.OracleGraph(a)
    pool = np.arange(0, 400, 37, dtype=np.int32)
    seen, todo = set(pool.tolist()), list(pool.tolist())
    while todo:
        v = todo.pop()
        for u in a["adj0"][v]:
            if u >= 0 and int(u) not in seen:
                seen.add(int(u))
                todo.append(int(u))
    reach = np.array(sorted(seen), dtype=np.int32)
    assert len(reach) >= 50
    gt, _ = orc.bruteforce_f64(x[reach], q, 10)
    ids, ds, st = orc.nsg_search(og, q, 10, 400, pool, order=orc.SEQ)
    np.testing.assert_array_equal(ids, reach[gt])
    assert np.all(st[:, 1] == len(reach))  # every reachable node evaluated once
