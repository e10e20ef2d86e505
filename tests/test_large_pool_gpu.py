"""Pools above 512 entries (the large-pool kernel) against the CPU oracle.

SearchArraySimulation takes any L (src/annsearch.cpp:182-197); the
reference's experiments sweep L up to 1,600 (scripts/parallel_search.sh:5)
and 8,000-16,000 (scripts/search.sh:7).  The large-pool kernel keeps the
pool in LDS and merges by binary search and a shift of the moved entries;
its results must equal the oracle's exactly: ids, distance bits, hops and,
wherever the visited table forgot nothing, distance counts.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _index(base, adj, ep, metric=0):
    from scattersearch_amd import Index
    ix = Index(base, metric)
    ix.set_graph(adj, ep)
    return ix


def _assert_same(gpu, ora, what, K=None):
    ids, dists, st = gpu
    o_ids, o_d, o_ndc, o_hops = ora
    nq = ids.shape[0]
    Ks = np.broadcast_to(np.asarray(K if K is not None else o_ids.shape[1]), (nq,))
    for i in range(nq):
        k = int(Ks[i])
        assert np.array_equal(ids[i, :k], o_ids[i, :k]), f"{what}: ids differ on query {i}"
        assert np.array_equal(dists[i, :k].view(np.uint32), o_d[i, :k].view(np.uint32)), f"{what}: query {i}"
    assert np.array_equal(st["hops"], o_hops), f"{what}: hop counts differ"
    exact = st["lossy"] == 0
    assert np.array_equal(st["ndc"][exact], o_ndc[exact]), f"{what}: distance counts differ"
    assert (st["ndc"] >= o_ndc).all()
    assert (st["status"] == 0).all()


@pytest.mark.parametrize("L", [513, 1024, 2048, 4096])
def test_large_pools_equal_oracle(c0, L):
    import oracle
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    q = c0["queries"][:48]
    for K in (10, 100):
        print(f"large pool: L={L} K={K}, {q.shape[0]} queries", flush=True)
        gpu = ix.search(q, K, L)
        ora = oracle.search_batch(c0["base"], c0["adj"], c0["ep"], q, K, L)
        _assert_same(gpu, ora, f"L={L} K={K}")


def test_large_pool_with_a_small_visited_table(c0):
    """256 slots at L = 1024: buckets fill, forgotten ids come back as
    candidates (pool members among them, dropped by the merge); results
    are unchanged, only distance counts grow."""
    import oracle
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    q = c0["queries"][:32]
    for bucket in (8, 16):
        gpu = ix.search(q, 10, 1024, visited_slots=256, visited_bucket=bucket)
        assert gpu[2]["lossy"].sum() > 0
        ora = oracle.search_batch(c0["base"], c0["adj"], c0["ep"], q, 10, 1024)
        _assert_same(gpu, ora, f"256 slots, buckets of {bucket}")


def test_pool_larger_than_the_base():
    """L above the number of points (the pool never fills, so the tail filter
    keeps the seed pool's last distance for the whole search): equal to the
    oracle, and every returned entry is a distinct point at its distance."""
    import oracle
    from scattersearch_amd import data
    base, q = data.mixture(3000, 8, 64, 6, 0.3, seed=91)
    adj, ep, _ = oracle.cpu_graph(base, 0, 16, 16, 32, passes=1)
    ix = _index(base, adj, ep)
    gpu = ix.search(q, 20, 4000)
    ora = oracle.search_batch(base, adj, ep, q, 20, 4000)
    _assert_same(gpu, ora, "L=4000 on 3000 points")
    ids, ds = gpu[0], gpu[1]
    for i in range(q.shape[0]):
        row = ids[i][ids[i] != 0xFFFFFFFF]
        assert len(set(row.tolist())) == row.size
        assert (np.diff(ds[i][: row.size]) >= 0).all()


def test_largest_pool(c0):
    """L = 16,384 (the reference's largest swept pool) on the 20k-point
    workload."""
    import oracle
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    q = c0["queries"][:4]
    print("large pool: L=16384, 4 queries", flush=True)
    gpu = ix.search(q, 10, 16384)
    ora = oracle.search_batch(c0["base"], c0["adj"], c0["ep"], q, 10, 16384)
    _assert_same(gpu, ora, "L=16384")


@pytest.mark.parametrize("dim,metric", [(200, 1), (40, 0), (96, 0)])
def test_large_pool_dims_and_metrics(dim, metric):
    import oracle
    from scattersearch_amd import data
    base, q = data.mixture(6000, 24, dim, 12, 0.3, seed=dim, lognormal_sigma=0.5 if metric else 0.0,
                           unit_norm=bool(metric))
    adj, ep, _ = oracle.cpu_graph(base, metric, 24, 24, 32, passes=1)
    ix = _index(base, adj, ep, metric)
    gpu = ix.search(q, 10, 777)
    ora = oracle.search_batch(base, adj, ep, q, 10, 777, metric)
    _assert_same(gpu, ora, f"d={dim} metric={metric}")


def test_mixed_k_and_l_across_512_in_one_launch(small):
    """Per-query K and L with some pools above 512: the whole launch runs the
    large-pool kernel; every query equals the oracle."""
    import oracle
    rng = np.random.default_rng(29)
    q = small["queries"]
    nq = q.shape[0]
    K = rng.choice([1, 10, 100], size=nq).astype(np.uint32)
    L = np.array([rng.integers(max(int(k), 16), 1500) for k in K], dtype=np.uint32)
    assert (L > 512).any() and (L <= 512).any()
    ix = _index(small["base"], small["adj"], small["ep"])
    gpu = ix.search(q, K, L, k_stride=100)
    ora = oracle.search_batch(small["base"], small["adj"], small["ep"], q, K, L, k_stride=100)
    _assert_same(gpu, ora, "mixed K/L", K=K)
    for i in range(nq):
        assert (gpu[0][i, K[i]:] == 0xFFFFFFFF).all()


def test_large_pool_launch_modes(small):
    """Per-query launches, hipGraph replays and the locality order give the
    same results as the batched launch."""
    from scattersearch_amd import SS_MODE_PER_QUERY
    from scattersearch_amd._lib import SS_ORDER_LOCALITY
    ix = _index(small["base"], small["adj"], small["ep"])
    q = small["queries"][:40]
    ref = ix.search(q, 10, 900)
    for kw in ({"mode": SS_MODE_PER_QUERY}, {"use_graph": True}, {"use_graph": True},
               {"query_order": SS_ORDER_LOCALITY}, {"blocks": 3}):
        got = ix.search(q, 10, 900, **kw)
        assert np.array_equal(got[0], ref[0]), kw
        assert np.array_equal(got[1].view(np.uint32), ref[1].view(np.uint32)), kw
        assert np.array_equal(got[2]["hops"], ref[2]["hops"]), kw


def test_large_pool_without_allocation(small):
    """ss_search_device_noalloc takes the large-pool kernel too."""
    import torch
    from scattersearch_amd import dim_padded
    from scattersearch_amd._lib import SearchParams, lib
    ix = _index(small["base"], small["adj"], small["ep"])
    q = small["queries"][:40]
    ref = ix.search(q, 10, 700)
    dq = torch.zeros((q.shape[0], dim_padded(q.shape[1])), dtype=torch.float32, device="cuda")
    dq[:, :q.shape[1]] = torch.from_numpy(q).cuda()
    ids = torch.zeros((q.shape[0], 10), dtype=torch.int32, device="cuda")
    ds = torch.zeros((q.shape[0], 10), dtype=torch.float32, device="cuda")
    p = SearchParams()
    vp = ctypes.c_void_p
    rc = lib.ss_search_device_noalloc(ix.handle, vp(dq.data_ptr()), q.shape[0], None, None, 10, 700, 10,
                                      ctypes.byref(p), vp(ids.data_ptr()), vp(ds.data_ptr()), None, None)
    torch.cuda.synchronize()
    assert rc == 0, lib.ss_last_error()
    assert np.array_equal(ids.cpu().numpy().view(np.uint32), ref[0])
    assert np.array_equal(ds.cpu().numpy().view(np.uint32), ref[1].view(np.uint32))


def test_large_pool_limits(small):
    from scattersearch_amd import SSError
    from scattersearch_amd._lib import SS_EINVAL, SS_MAX_L_LARGE
    ix = _index(small["base"], small["adj"], small["ep"])
    q = small["queries"][:2]
    for L, kw in ((SS_MAX_L_LARGE + 1, {}), (1024, {"expand_width": 2}), (1024, {"waves_per_query": 2})):
        with pytest.raises(SSError) as e:
            ix.search(q, 10, L, **kw)
        assert e.value.code == SS_EINVAL, (L, kw)


@pytest.mark.parametrize("L", [16, 48, 64, 200, 512])
def test_shift_layout_at_small_pools_equals_oracle(c0, L):
    """The shift-merge kernel asked for at pools of 512 entries or fewer
    (pool_layout SS_POOL_SHIFT) gives the oracle's results, like the kernels
    it replaces there."""
    import oracle
    from scattersearch_amd._lib import SS_POOL_SHIFT
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    q = c0["queries"][:64]
    gpu = ix.search(q, 10, L, pool_layout=SS_POOL_SHIFT)
    ora = oracle.search_batch(c0["base"], c0["adj"], c0["ep"], q, 10, L)
    _assert_same(gpu, ora, f"shift layout L={L}")
    ref = ix.search(q, 10, L)
    assert np.array_equal(gpu[0], ref[0]) and np.array_equal(gpu[2]["hops"], ref[2]["hops"])


def test_shift_layout_in_pool_size_classes(small):
    """Per-query L split into pool-size classes, every class on the
    shift-merge kernel (each takes only its own range of the order)."""
    import oracle
    from scattersearch_amd._lib import SS_POOL_SHIFT, SS_CLASSES_AUTO, SS_CLASSES_SERIAL
    rng = np.random.default_rng(41)
    q = small["queries"]
    nq = q.shape[0]
    K = rng.choice([1, 10, 32], size=nq).astype(np.uint32)
    L = np.array([rng.integers(max(int(k), 16), 513) for k in K], dtype=np.uint32)
    ix = _index(small["base"], small["adj"], small["ep"])
    ora = oracle.search_batch(small["base"], small["adj"], small["ep"], q, K, L, k_stride=32)
    for classes in (SS_CLASSES_AUTO, SS_CLASSES_SERIAL):
        gpu = ix.search(q, K, L, k_stride=32, pool_layout=SS_POOL_SHIFT, pool_classes=classes, use_graph=True)
        _assert_same(gpu, ora, f"shift layout, classes {classes}", K=K)
