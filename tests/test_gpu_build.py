"""GPU HNSW construction (gvs_hnsw_build_device): hnswlib's addPoint run in
batches on the device.  Pinned three ways: batches of one in hnswlib's
scalar order give the reference hnswlib graph (the golden .bin indexes,
written by the reference's own hnswlib); larger batches in that order give
the host batched builder's graph bit for bit; the fast order gives a sound
graph whose recall matches the CPU-built graph's."""
import numpy as np
import pytest

import gvs
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def same_graph(a, b):
    assert a["entry"] == b["entry"] and a["max_level"] == b["max_level"]
    np.testing.assert_array_equal(a["adj0"], b["adj0"])
    np.testing.assert_array_equal(a["upper_off"], b["upper_off"])
    np.testing.assert_array_equal(a["adj_up"], b["adj_up"])


@pytest.mark.parametrize("name", ["l2_n1000_d16_M8", "ip_n1000_d24_M6", "l2_n400_d8_M4"])
def test_gpu_build_batches_of_one_is_the_reference_graph(name):
    import test_oracle as T
    z, g = T.load_case(name)
    print(f"case gpu build {name} batches of one", flush=True)
    mine = gvs.HostGraph.build_hnsw_gpu(z["x"], M=int(z["M"]), ef_construction=int(z["efc"]), seed=100,
                                        metric=int(z["metric"]), exact_order=True, batch_ratio=-1)
    same_graph(mine.arrays(), g.arrays())


@pytest.mark.parametrize("metric,ratio", [(gvs.L2, 0.05), (gvs.IP, 0.1), (gvs.L2, 0.5)])
def test_gpu_build_batched_equals_host_batched(metric, ratio):
    """Batches of ratio x the graph (up to one half: many new links per
    target, re-pruned at once) in hnswlib's order: the GPU graph equals the
    host batched builder's."""
    x = gvs.gen_mixture(5000, 48, 20, 81, 82, normalize=metric == gvs.IP)
    print(f"case gpu build batched metric={metric} ratio={ratio}", flush=True)
    a = gvs.HostGraph.build_hnsw_gpu(x, M=10, ef_construction=60, metric=metric, exact_order=True,
                                     batch_ratio=ratio)
    b = gvs.HostGraph.build_hnsw_batched(x, M=10, ef_construction=60, metric=metric, nthreads=8,
                                         exact_order=True, batch_ratio=ratio)
    same_graph(a.arrays(), b.arrays())


def test_gpu_build_fast_order_sound_deterministic_and_as_good():
    x = gvs.gen_mixture(30000, 128, 32, 91, 92)
    q = gvs.gen_mixture(1000, 128, 32, 91, 93)
    g1, st = gvs.HostGraph.build_hnsw_gpu(x, M=16, ef_construction=100, with_stats=True)
    print(f"case gpu build fast 30k x 128: {st}", flush=True)
    g2 = gvs.HostGraph.build_hnsw_gpu(x, M=16, ef_construction=100)
    same_graph(g1.arrays(), g2.arrays())
    a = g1.arrays()
    adj = a["adj0"]
    valid = adj >= 0
    assert np.all(np.diff(valid.astype(np.int8), axis=1) <= 0)  # rows compact
    assert not np.any(adj == np.arange(len(adj))[:, None])
    gc = gvs.HostGraph.build_hnsw(x, M=16, ef_construction=100, nthreads=16)
    gt, _ = orc.bruteforce_f64(x, q, 10)
    rg = orc.recall_at(gvs.DeviceIndex(g1).search(q, k=10, ef=32)[0], gt, 10)
    rc = orc.recall_at(gvs.DeviceIndex(gc).search(q, k=10, ef=32)[0], gt, 10)
    print(f"  recall@10 at ef 32: GPU-built {rg:.4f}, CPU-built {rc:.4f}", flush=True)
    assert rg >= rc - 0.01
