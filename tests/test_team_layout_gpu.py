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


@pytest.mark.parametrize("dim,metric", [(200, 1), (40, 0), (96, 0)])
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
               {"team_size": 8, "waves_per_query": 2}, {"pool_layout": 3}):
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
