"""Distance team size (4 or 8 lanes per distance) and pool layout (pool read
into registers once per step, or scanned in LDS and re-read only when a
candidate is admitted) are launch choices of the autotuner: neither may
change a result.  Every combination must return the oracle's ids, distance
bits, NDC and hops (SearchArraySimulation, src/annsearch.cpp:182-259; the
distance routines src/distance.cpp:19-56, 115-149), for 128-d L2, 200-d inner
product (leading tail block) and a runtime-size dimension, across pool sizes
whose bodies differ (one register chunk, 2, 4 and 8 chunks)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _index(base, adj, ep, metric=0):
    from scattersearch_amd import Index
    ix = Index(base, metric)
    ix.set_graph(adj, ep)
    return ix


def _same(gpu, ora, what):
    ids, ds, st = gpu
    o_ids, o_d, o_ndc, o_hops = ora
    k = o_ids.shape[1]
    assert np.array_equal(ids[:, :k], o_ids), f"{what}: ids"
    assert np.array_equal(ds[:, :k].view(np.uint32), o_d.view(np.uint32)), f"{what}: distance bits"
    assert np.array_equal(st["hops"], o_hops), f"{what}: hops"
    clean = st["lossy"] == 0
    assert np.array_equal(st["ndc"][clean], o_ndc[clean]), f"{what}: ndc"


@pytest.mark.parametrize("team", [4, 8])
def test_teams_and_layouts_equal_oracle_c0(c0, team):
    import oracle
    from scattersearch_amd._lib import SS_POOL_AUTO, SS_POOL_LDS, SS_POOL_REGISTERS
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    q = c0["queries"][:120]
    for L in (16, 48, 64, 128, 256, 512):
        ora = oracle.search_batch(c0["base"], c0["adj"], c0["ep"], q, 10, L)
        layouts = [SS_POOL_AUTO, SS_POOL_LDS] + ([SS_POOL_REGISTERS] if L <= 64 else [])
        for layout in layouts:
            gpu = ix.search(q, 10, L, team_size=team, pool_layout=layout)
            _same(gpu, ora, f"team {team} layout {layout} L={L}")


@pytest.mark.parametrize("dim,metric", [(200, 1), (40, 0), (96, 0), (8, 0), (8, 1), (20, 1), (24, 0), (72, 1),
                                        (136, 0), (520, 1), (1000, 0)])
def test_team8_equals_oracle_other_dims(dim, metric):
    import oracle
    from scattersearch_amd import data
    from scattersearch_amd.graph import medoid
    from scattersearch_amd._lib import SS_POOL_LDS
    base, q = data.mixture(3000, 80, dim, 12, 0.3, seed=70 + dim, lognormal_sigma=0.5 * metric,
                           unit_norm=bool(metric))
    adj, ep, _ = oracle.cpu_graph(base, metric, 24, 24, 32, passes=1)
    ix = _index(base, adj, ep, metric)
    rng = np.random.default_rng(dim)
    K = rng.choice([1, 10, 32], size=q.shape[0]).astype(np.uint32)
    L = np.array([rng.integers(max(int(k), 16), 300) for k in K], dtype=np.uint32)
    ora = oracle.search_batch(base, adj, ep, q, K, L, metric, k_stride=32)
    for team in (4, 8):
        for layout in (0, SS_POOL_LDS):
            gpu = ix.search(q, K, L, k_stride=32, team_size=team, pool_layout=layout)
            ids, ds, st = gpu
            for i in range(q.shape[0]):
                k = int(K[i])
                assert np.array_equal(ids[i, :k], ora[0][i, :k]), (dim, team, layout, i)
                assert np.array_equal(ds[i, :k].view(np.uint32), ora[1][i, :k].view(np.uint32)), (dim, team, i)
            assert np.array_equal(st["hops"], ora[3]), (dim, team, layout)


def test_team8_batched_locality_order_and_graph_replay(c0):
    """Team 8 in the batched path: a locality-ordered batch of 5,000 queries,
    captured once and replayed, equals the four-lane search bit for bit."""
    base = c0["base"]
    rng = np.random.default_rng(3)
    q = base[rng.choice(base.shape[0], 5000, replace=False)] + np.float32(0.05) * rng.standard_normal(
        (5000, base.shape[1]), dtype=np.float32)
    ix = _index(base, c0["adj"], c0["ep"])
    ref = ix.search(q, 10, 48)
    for rep in range(3):
        got = ix.search(q, 10, 48, team_size=8, use_graph=True)
        assert np.array_equal(got[0], ref[0]), rep
        assert np.array_equal(got[1].view(np.uint32), ref[1].view(np.uint32)), rep
        assert np.array_equal(got[2]["hops"], ref[2]["hops"]), rep


def test_bad_team_and_layout_refused(small):
    from scattersearch_amd import SSError
    from scattersearch_amd._lib import SS_EINVAL, SS_POOL_REGISTERS
    ix = _index(small["base"], small["adj"], small["ep"])
    q = small["queries"][:4]
    for kw in ({"team_size": 2}, {"team_size": 16}, {"team_size": 8, "expand_width": 2},
               {"team_size": 8, "waves_per_query": 2}, {"pool_layout": 4}):
        with pytest.raises(SSError) as e:
            ix.search(q, 10, 48, **kw)
        assert e.value.code == SS_EINVAL, kw
    with pytest.raises(SSError) as e:
        ix.search(q, 10, 128, pool_layout=SS_POOL_REGISTERS)
    assert e.value.code == SS_EINVAL


def test_autotune_reports_team_and_layout(c0):
    import oracle
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    q = c0["queries"]
    for L in (48, 128):
        t = ix.autotune(q, 10, L, max_recall_loss=0.0)
        assert t["team_size"] in (4, 8) and t["pool_layout"] in (1, 2), t
        assert t["expand_width"] == 1 and t["waves_per_query"] == 1
        ora = oracle.search_batch(c0["base"], c0["adj"], c0["ep"], q, 10, L)
        _same(ix.search(q, 10, L), ora, f"tuned L={L}")


def test_pool_classes_equal_oracle_and_each_other(c0):
    """A launch mixing K and L up to 512 is split by pool size (L <= 64,
    <= 128, <= 256, <= 512) into kernels with their own LDS shapes, run side
    by side, one after another, or not split: ids, distance bits, hops and
    NDC equal the oracle in every mode, also when the launch is captured in
    a graph and replayed, and through the no-allocation entry point."""
    import ctypes
    import torch
    import oracle
    from scattersearch_amd import dim_padded
    from scattersearch_amd._lib import SS_CLASSES_AUTO, SS_CLASSES_OFF, SS_CLASSES_SERIAL, SearchParams, lib
    base, adj, ep = c0["base"], c0["adj"], c0["ep"]
    rng = np.random.default_rng(12)
    nq = 600
    q = base[rng.choice(base.shape[0], nq, replace=False)] + np.float32(0.05) * rng.standard_normal(
        (nq, base.shape[1]), dtype=np.float32)
    K = rng.choice([1, 10, 100], size=nq).astype(np.uint32)
    L = np.array([rng.integers(int(k), 513) for k in K], dtype=np.uint32)
    ora = oracle.search_batch(base, adj, ep, q, K, L, k_stride=100)
    ix = _index(base, adj, ep)

    def check(ids, ds, st, what):
        for i in range(nq):
            k = int(K[i])
            assert np.array_equal(ids[i, :k], ora[0][i, :k]), (what, i)
            assert np.array_equal(ds[i, :k].view(np.uint32), ora[1][i, :k].view(np.uint32)), (what, i)
        assert np.array_equal(st["hops"], ora[3]), what
        clean = st["lossy"] == 0
        assert np.array_equal(st["ndc"][clean], ora[2][clean]), what

    from scattersearch_amd._lib import SS_ORDER_INDEX, SS_ORDER_LOCALITY
    for mode in (SS_CLASSES_AUTO, SS_CLASSES_SERIAL, SS_CLASSES_OFF):
        for rep in range(3 if mode == SS_CLASSES_AUTO else 1):
            ids, ds, st = ix.search(q, K, L, k_stride=100, pool_classes=mode, use_graph=rep > 0)
            check(ids, ds, st, f"mode {mode} rep {rep}")
    # the classes taken longest pool first (the default), in no order, and
    # by locality
    for order in (SS_ORDER_INDEX, SS_ORDER_LOCALITY):
        ids, ds, st = ix.search(q, K, L, k_stride=100, query_order=order)
        check(ids, ds, st, f"order {order}")
    # the no-allocation entry point on device buffers
    dp = dim_padded(base.shape[1])
    dq = torch.zeros((nq, dp), dtype=torch.float32, device="cuda")
    dq[:, :base.shape[1]] = torch.from_numpy(q).cuda()
    dK = torch.from_numpy(K.astype(np.int32)).cuda()
    dL = torch.from_numpy(L.astype(np.int32)).cuda()
    out = torch.zeros((nq, 100), dtype=torch.int32, device="cuda")
    dist = torch.zeros((nq, 100), dtype=torch.float32, device="cuda")
    stt = torch.zeros((nq, 8), dtype=torch.int32, device="cuda")
    vp = ctypes.c_void_p
    p = SearchParams()
    rc = lib.ss_search_device_noalloc(ix.handle, vp(dq.data_ptr()), nq, vp(dK.data_ptr()), vp(dL.data_ptr()), 100,
                                      512, 100, ctypes.byref(p), vp(out.data_ptr()), vp(dist.data_ptr()),
                                      vp(stt.data_ptr()), None)
    torch.cuda.synchronize()
    assert rc == 0, lib.ss_last_error()
    from scattersearch_amd.index import STATS_DTYPE
    st = stt.cpu().numpy().view(np.uint32).reshape(-1).view(STATS_DTYPE)
    check(out.cpu().numpy().view(np.uint32), dist.cpu().numpy(), st, "no-alloc")
