"""The locality order of batched searches (SS_ORDER_LOCALITY, the default for
batches of at least SS_ORDER_MIN_BATCH queries) changes only which wavefront
takes which query and when.  Ids, distances (bit for bit), distance counts,
hop counts and lossy-insert counts must equal the index-order search, and
the oracle's (SearchArraySimulation, src/annsearch.cpp:182-259).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    assert np.array_equal(a[0], b[0]), f"{what}: ids differ"
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32)), f"{what}: distances differ"
    for f in ("ndc", "hops", "lossy", "status"):
        assert np.array_equal(a[2][f], b[2][f]), f"{what}: {f} differs"


@pytest.fixture(scope="module")
def big():
    """50k x 128 mixture, 9,000 queries (above SS_ORDER_MIN_BATCH), graph from
    the GPU builder."""
    from scattersearch_amd import Index, data
    from scattersearch_amd.graph import gpu_build_graph
    base, queries = data.mixture(50_000, 9_000, 128, 100, 0.25, seed=4242)
    ix = Index(base, 0)
    adj, ep, _ = gpu_build_graph(ix, base, 32, 32, 48)
    return dict(ix=ix, base=base, queries=queries, adj=adj, ep=ep)


def test_default_order_of_a_large_batch_is_exact(big):
    import oracle
    from scattersearch_amd._lib import SS_ORDER_INDEX, SS_ORDER_MIN_BATCH
    assert big["queries"].shape[0] >= SS_ORDER_MIN_BATCH
    ix = big["ix"]
    loc = ix.search(big["queries"], 10, 48)
    idx = ix.search(big["queries"], 10, 48, query_order=SS_ORDER_INDEX)
    _same(loc, idx, "locality vs index order")
    o_ids, o_d, o_ndc, o_hops = oracle.search_batch(big["base"], big["adj"], big["ep"], big["queries"], 10, 48)
    assert np.array_equal(loc[0], o_ids)
    assert np.array_equal(loc[1].view(np.uint32), o_d.view(np.uint32))
    assert np.array_equal(loc[2]["hops"], o_hops)
    exact = loc[2]["lossy"] == 0
    assert np.array_equal(loc[2]["ndc"][exact], o_ndc[exact])


def test_locality_order_with_per_query_k_and_l(big):
    from scattersearch_amd._lib import SS_ORDER_INDEX, SS_ORDER_LOCALITY
    rng = np.random.default_rng(5)
    nq = 5000
    K = rng.choice([1, 10, 100], size=nq).astype(np.uint32)
    L = np.maximum(K, rng.integers(16, 200, size=nq)).astype(np.uint32)
    q = big["queries"][:nq]
    a = big["ix"].search(q, K, L, query_order=SS_ORDER_LOCALITY)
    b = big["ix"].search(q, K, L, query_order=SS_ORDER_INDEX)
    _same(a, b, "per-query K and L")


def test_longest_first_order_of_one_kernel(big):
    """A batch of at least SS_ORDER_MIN_BATCH queries with per-query L above
    64 searched by one kernel (pool-size classes off, and the large-pool
    kernel) takes its queries longest pool first: same results as the index
    order and, on a sample, as the oracle."""
    import oracle
    from scattersearch_amd._lib import SS_CLASSES_OFF, SS_ORDER_INDEX, SS_ORDER_MIN_BATCH
    rng = np.random.default_rng(9)
    nq = SS_ORDER_MIN_BATCH + 700
    q = big["queries"][:nq]
    for lo, hi in ((1, 257), (400, 700)):
        K = rng.choice([1, 10, 100], size=nq).astype(np.uint32)
        L = np.maximum(K, rng.integers(lo, hi, size=nq)).astype(np.uint32)
        print(f"longest-first, L in {lo}..{hi - 1}, {nq} queries", flush=True)
        a = big["ix"].search(q, K, L, k_stride=100, pool_classes=SS_CLASSES_OFF)
        b = big["ix"].search(q, K, L, k_stride=100, pool_classes=SS_CLASSES_OFF, query_order=SS_ORDER_INDEX)
        _same(a, b, f"one kernel, L {lo}..{hi - 1}")
        sel = rng.choice(nq, 120, replace=False)
        o_ids, o_d, o_ndc, o_hops = oracle.search_batch(big["base"], big["adj"], big["ep"], q[sel], K[sel], L[sel],
                                                        k_stride=100)
        for j, i in enumerate(sel):
            k = int(K[i])
            assert np.array_equal(a[0][i, :k], o_ids[j, :k]), (lo, i)
            assert np.array_equal(a[1][i, :k].view(np.uint32), o_d[j, :k].view(np.uint32)), (lo, i)
        assert np.array_equal(a[2]["hops"][sel], o_hops)


def test_longest_first_order_without_allocation(big):
    """The longest-first order of one kernel also runs on the no-allocation
    entry point (keys and order in its scratch array): same results as
    ss_search_device, for both pool kernels."""
    import ctypes
    import torch
    from scattersearch_amd import dim_padded
    from scattersearch_amd._lib import SS_CLASSES_OFF, SS_ORDER_MIN_BATCH, SearchParams, lib
    ix, q = big["ix"], big["queries"]
    nq = SS_ORDER_MIN_BATCH + 500
    rng = np.random.default_rng(17)
    dp = dim_padded(q.shape[1])
    dq = torch.zeros((nq, dp), dtype=torch.float32, device="cuda")
    dq[:, :q.shape[1]] = torch.from_numpy(q[:nq]).cuda()
    vp = ctypes.c_void_p
    for lo, hi in ((65, 300), (513, 700)):
        K = rng.choice([1, 10, 100], size=nq).astype(np.int32)
        L = np.maximum(K, rng.integers(lo, hi + 1, size=nq)).astype(np.int32)
        dK = torch.from_numpy(K).cuda()
        dL = torch.from_numpy(L).cuda()
        out = []
        for fn in (lib.ss_search_device, lib.ss_search_device_noalloc):
            print(f"longest-first without allocation: L {lo}..{hi}, {fn.__name__}", flush=True)
            p = SearchParams()
            p.pool_classes = SS_CLASSES_OFF
            ids = torch.full((nq, 100), -1, dtype=torch.int32, device="cuda")
            ds = torch.zeros((nq, 100), dtype=torch.float32, device="cuda")
            st = torch.zeros((nq, 8), dtype=torch.int32, device="cuda")
            rc = fn(ix.handle, vp(dq.data_ptr()), nq, vp(dK.data_ptr()), vp(dL.data_ptr()), 100, hi, 100,
                    ctypes.byref(p), vp(ids.data_ptr()), vp(ds.data_ptr()), vp(st.data_ptr()), None)
            torch.cuda.synchronize()
            assert rc == 0, lib.ss_last_error()
            out.append((ids.cpu().numpy(), ds.cpu().numpy().view(np.uint32), st.cpu().numpy()))
        (ia, da, sa), (ib, db, sb) = out
        for i in range(nq):
            k = int(K[i])
            assert np.array_equal(ia[i, :k], ib[i, :k]), (lo, i)
            assert np.array_equal(da[i, :k], db[i, :k]), (lo, i)
        assert np.array_equal(sa[:, :4], sb[:, :4]), lo  # ndc, hops, lossy, status


def test_locality_order_graph_replays(big):
    """hipGraph replays of the ordered launch (keys, sort, search), with eager
    work in between, each equal to the eager search."""
    from scattersearch_amd._lib import SS_ORDER_LOCALITY
    ix = big["ix"]
    eager = ix.search(big["queries"], 10, 64, query_order=SS_ORDER_LOCALITY)
    for r in range(3):
        g = ix.search(big["queries"], 10, 64, query_order=SS_ORDER_LOCALITY, use_graph=True)
        _same(eager, g, f"replay {r}")
        mid = ix.search(big["queries"][:50], 5, 20)
        assert mid[0].shape == (50, 5)


def test_orders_of_a_batch_sorted_in_several_passes(small):
    """Batches above 16,384 queries are sorted by the order kernel in several
    passes over the keys: 40,000 distinct queries in the locality order, per
    query K and L split into pool-size classes, and one kernel longest pool
    first all equal the index order, and a sample equals the oracle."""
    import oracle
    from scattersearch_amd import Index
    from scattersearch_amd._lib import SS_CLASSES_OFF, SS_ORDER_INDEX, SS_ORDER_LOCALITY
    rng = np.random.default_rng(40)
    ix = Index(small["base"], 0)
    ix.set_graph(small["adj"], small["ep"])
    nq = 40_000
    base_q = small["queries"]
    q = (base_q[rng.integers(0, base_q.shape[0], size=nq)]
         + rng.normal(0, 0.05, size=(nq, base_q.shape[1]))).astype(np.float32)
    print(f"locality order, {nq} queries", flush=True)
    a = ix.search(q, 10, 48, query_order=SS_ORDER_LOCALITY)
    b = ix.search(q, 10, 48, query_order=SS_ORDER_INDEX)
    _same(a, b, "locality vs index order, 40,000 queries")
    sel = rng.choice(nq, 300, replace=False)
    o_ids, o_d, o_ndc, o_hops = oracle.search_batch(small["base"], small["adj"], small["ep"], q[sel], 10, 48)
    assert np.array_equal(a[0][sel], o_ids)
    assert np.array_equal(a[1][sel].view(np.uint32), o_d.view(np.uint32))
    assert np.array_equal(a[2]["hops"][sel], o_hops)
    K = rng.choice([1, 10], size=nq).astype(np.uint32)
    L = np.maximum(K, rng.integers(8, 300, size=nq)).astype(np.uint32)
    print("pool-size classes, 40,000 queries", flush=True)
    c = ix.search(q, K, L, k_stride=10)
    d = ix.search(q, K, L, k_stride=10, query_order=SS_ORDER_INDEX)
    _same(c, d, "classes vs index order")
    L2 = np.maximum(K, rng.integers(65, 300, size=nq)).astype(np.uint32)
    print("longest pool first, 40,000 queries", flush=True)
    e = ix.search(q, K, L2, k_stride=10, pool_classes=SS_CLASSES_OFF)
    f = ix.search(q, K, L2, k_stride=10, pool_classes=SS_CLASSES_OFF, query_order=SS_ORDER_INDEX)
    _same(e, f, "longest first vs index order")


@pytest.mark.parametrize("nq", [1, 3, 7, 8, 9, 64, 1000])
def test_forced_locality_on_small_batches(small, nq):
    """Small and odd batch sizes, one query up to less than the default threshold."""
    from scattersearch_amd import Index
    from scattersearch_amd._lib import SS_ORDER_INDEX, SS_ORDER_LOCALITY
    ix = Index(small["base"], 0)
    ix.set_graph(small["adj"], small["ep"])
    q = np.resize(small["queries"], (nq, small["queries"].shape[1]))
    a = ix.search(q, 10, 48, query_order=SS_ORDER_LOCALITY)
    b = ix.search(q, 10, 48, query_order=SS_ORDER_INDEX)
    _same(a, b, f"nq={nq}")


@pytest.mark.parametrize("dim", [8, 100, 200])
@pytest.mark.parametrize("metric", [0, 1])
def test_forced_locality_dims_and_metrics(dim, metric):
    """Keys use the first min(dim_pad, 32) coordinates."""
    import oracle
    from scattersearch_amd import Index, data
    from scattersearch_amd._lib import SS_ORDER_LOCALITY
    base, q = data.mixture(2000, 300, dim, 6, 0.3, seed=dim + 7 * metric,
                           lognormal_sigma=0.5 if metric else 0.0, unit_norm=bool(metric))
    adj, ep, _ = oracle.cpu_graph(base, metric, 20, 20, 24, passes=1)
    ix = Index(base, metric)
    ix.set_graph(adj, ep)
    gpu = ix.search(q, 10, 40, visited_slots=8192, query_order=SS_ORDER_LOCALITY)
    o_ids, o_d, o_ndc, o_hops = oracle.search_batch(base, adj, ep, q, 10, 40, metric=metric)
    assert gpu[2]["lossy"].sum() == 0
    assert np.array_equal(gpu[0], o_ids)
    assert np.array_equal(gpu[1].view(np.uint32), o_d.view(np.uint32))
    assert np.array_equal(gpu[2]["ndc"], o_ndc)
    assert np.array_equal(gpu[2]["hops"], o_hops)


def test_unknown_query_order_rejected(small):
    from scattersearch_amd import Index, SSError
    ix = Index(small["base"], 0)
    ix.set_graph(small["adj"], small["ep"])
    with pytest.raises(SSError):
        ix.search(small["queries"], 10, 48, query_order=3)
