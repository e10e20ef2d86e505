"""GPU parity of the search kernel against the CPU oracle
(SearchArraySimulation, src/annsearch.cpp:182-259).

The kernel computes distances with the reference's own arithmetic and
reproduces the sequential pool semantics exactly, so on the same graph,
entry point, L and K the ids, the distances (bit for bit), the distance
counts and the hop counts must all equal the oracle's.
"""
import numpy as np
import pytest

from conftest import random_graph

pytestmark = pytest.mark.gpu

NONE = 0xFFFFFFFF


def _index(base, adj, ep, metric=0):
    from scattersearch_amd import Index
    ix = Index(base, metric)
    ix.set_graph(adj, ep)
    return ix


def _assert_same(gpu, ora, check_ndc=True, what=""):
    ids, dists, st = gpu
    o_ids, o_d, o_ndc, o_hops = ora
    k = o_ids.shape[1]
    bad = np.nonzero((ids[:, :k] != o_ids).any(1))[0]
    assert bad.size == 0, f"{what}: ids differ on queries {bad[:10]}"
    db = dists[:, :k].view(np.uint32) != o_d.view(np.uint32)
    assert not db.any(), f"{what}: distances differ on queries {np.nonzero(db.any(1))[0][:10]}"
    assert np.array_equal(st["hops"], o_hops), f"{what}: hop counts differ"
    if check_ndc:
        assert np.array_equal(st["ndc"], o_ndc), f"{what}: distance counts differ"


def test_c0_parity_bit_exact(c0):
    import oracle
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    gpu = ix.search(c0["queries"], 10, 64, visited_slots=16384)
    assert gpu[2]["lossy"].sum() == 0
    ora = oracle.search_batch(c0["base"], c0["adj"], c0["ep"], c0["queries"], 10, 64)
    _assert_same(gpu, ora, what="c0 L=64")


@pytest.mark.parametrize("L", [10, 32, 128, 256, 512])
def test_c0_pool_sizes(c0, L):
    import oracle
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    K = min(10, L)
    gpu = ix.search(c0["queries"][:100], K, L, visited_slots=32768 if L >= 256 else 16384)
    ora = oracle.search_batch(c0["base"], c0["adj"], c0["ep"], c0["queries"][:100], K, L)
    _assert_same(gpu, ora, check_ndc=gpu[2]["lossy"].sum() == 0, what=f"L={L}")


def test_entry_row_repeats_kept_other_rows_deduplicated(small):
    """Repeats in the entry point's row seed the pool as in the reference
    (two entries, src/annsearch.cpp:187-214); repeats elsewhere are dropped at
    load (the reference's visited set skips them).  Results, distances, NDC
    and hops equal the oracle on the graph as given."""
    import oracle
    base, q = small["base"], small["queries"]
    adj = small["adj"].copy()
    ep = small["ep"]
    deg = int((adj[ep] != NONE).sum())
    assert deg >= 4
    adj[ep, 1] = adj[ep, 0]  # the entry row names its first neighbour twice
    other = (ep + 1) % adj.shape[0]
    dother = int((adj[other] != NONE).sum())
    adj[other, dother - 1] = adj[other, 0]
    ix = _index(base, adj, ep)
    got, _ = ix.get_graph()
    assert got[ep].tolist() == adj[ep].tolist()
    assert (got[other] != NONE).sum() == dother - 1
    for L in (deg, 48):
        gpu = ix.search(q, 10, L, visited_slots=8192)
        ora = oracle.search_batch(base, adj, ep, q, 10, L)
        _assert_same(gpu, ora, what=f"repeated entry-row id, L={L}")


def test_small_pool_with_16bit_visited_slots_equals_oracle():
    """200-d inner product at L = 48 with 8,192 visited slots: the 32-bit
    table would allow 4 wavefronts per CU where the registers allow 8, so
    the one-read register pool runs with the 16-bit table; results equal the
    oracle (ids, distance bits, NDC, hops)."""
    import oracle
    from scattersearch_amd import data
    from scattersearch_amd.graph import medoid
    base, q = data.mixture(4000, 100, 200, 16, 0.25, seed=61, lognormal_sigma=0.5, unit_norm=True)
    knn = oracle.bruteforce(base, base, 24, 1, exclude=np.arange(4000, dtype=np.uint32))[0]
    ep = medoid(base)
    ix = _index(base, knn, ep, metric=1)
    for K, L in ((10, 48), (1, 20), (48, 64)):
        gpu = ix.search(q, K, L, visited_slots=8192)
        ora = oracle.search_batch(base, knn, ep, q, K, L, 1)
        _assert_same(gpu, ora, what=f"16-bit visited slots, K={K} L={L}")


def test_per_query_launch_equals_batched(small):
    ix = _index(small["base"], small["adj"], small["ep"])
    from scattersearch_amd import SS_MODE_PER_QUERY
    a = ix.search(small["queries"], 10, 48)
    b = ix.search(small["queries"], 10, 48, mode=SS_MODE_PER_QUERY)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    for f in ("ndc", "hops", "lossy"):
        assert np.array_equal(a[2][f], b[2][f])
    assert (a[2]["t_end"] >= a[2]["t_start"]).all()


@pytest.mark.parametrize("wpb", [1, 2, 4, 8])
def test_launch_shapes_same_results(small, wpb):
    ix = _index(small["base"], small["adj"], small["ep"])
    ref = ix.search(small["queries"], 10, 48, waves_per_block=1, blocks=1)
    out = ix.search(small["queries"], 10, 48, waves_per_block=wpb)
    assert np.array_equal(ref[0], out[0])
    assert np.array_equal(ref[2]["ndc"], out[2]["ndc"])


def test_grid_sizes_take_every_query_once(small):
    """Each wavefront takes its first query from its place in the grid (blocks
    numbered XCD by XCD) and later ones from the launch's counter: grids of
    1 to 1,000 blocks, not multiples of 8, larger and smaller than the batch,
    in index and locality order, for both pool kernels and for pool-size
    classes (each class a range of the order), all return the oracle's
    result for every query."""
    import oracle
    from scattersearch_amd._lib import SS_ORDER_INDEX, SS_ORDER_LOCALITY
    ix = _index(small["base"], small["adj"], small["ep"])
    q = np.resize(small["queries"], (150, small["queries"].shape[1]))
    rng = np.random.default_rng(21)
    Ku = rng.choice([1, 10], size=q.shape[0]).astype(np.uint32)
    Lu = np.maximum(Ku, rng.integers(8, 300, size=q.shape[0])).astype(np.uint32)
    want = {L: oracle.search_batch(small["base"], small["adj"], small["ep"], q, 10, L) for L in (48, 600)}
    want_u = oracle.search_batch(small["base"], small["adj"], small["ep"], q, Ku, Lu, k_stride=10)
    for blocks in (1, 3, 7, 9, 13, 100, 1000):
        for order in (SS_ORDER_INDEX, SS_ORDER_LOCALITY):
            for wpb in (1, 4):
                what = f"blocks={blocks} order={order} wpb={wpb}"
                print(what, flush=True)
                _assert_same(ix.search(q, 10, 48, blocks=blocks, query_order=order, waves_per_block=wpb),
                             want[48], what=what + " L=48")
            big = ix.search(q, 10, 600, blocks=blocks, query_order=order)
            _assert_same(big, want[600], check_ndc=False, what=f"blocks={blocks} order={order} L=600")
            exact = big[2]["lossy"] == 0  # the large-pool table may forget ids: only NDC grows
            assert np.array_equal(big[2]["ndc"][exact], want[600][2][exact]), (blocks, order)
            got = ix.search(q, Ku, Lu, k_stride=10, blocks=blocks, query_order=order)
            for i in range(q.shape[0]):
                k = int(Ku[i])
                assert np.array_equal(got[0][i, :k], want_u[0][i, :k]), (blocks, order, i)
            assert np.array_equal(got[2]["hops"], want_u[3]), (blocks, order)


def test_uneven_k_and_l_in_one_launch(small):
    import oracle
    rng = np.random.default_rng(3)
    nq = small["queries"].shape[0]
    K = rng.choice([1, 10, 50], size=nq).astype(np.uint32)
    L = np.array([rng.integers(k, 257) for k in K], dtype=np.uint32)
    ix = _index(small["base"], small["adj"], small["ep"])
    ids, d, st = ix.search(small["queries"], K, L, k_stride=64, visited_slots=16384)
    for q in range(nq):
        o_ids, o_d, o_ndc, o_hops = oracle.search(small["base"], small["adj"], small["ep"],
                                                  small["queries"][q], int(K[q]), int(L[q]))
        k = int(K[q])
        assert np.array_equal(ids[q, :k], o_ids), f"query {q} K={k} L={L[q]}"
        assert np.array_equal(d[q, :k].view(np.uint32), o_d.view(np.uint32))
        assert (ids[q, k:] == NONE).all() and np.isinf(d[q, k:]).all()
        assert st["hops"][q] == o_hops
        if st["lossy"][q] == 0:
            assert st["ndc"][q] == o_ndc


@pytest.mark.parametrize("slots", [256, 512, 1024])
def test_lossy_visited_hash_keeps_results(c0, slots):
    """A visited hash too small to remember everything only costs repeated
    distance computations: ids, distances and hops are unchanged."""
    import oracle
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    gpu = ix.search(c0["queries"], 10, 128, visited_slots=slots)
    ora = oracle.search_batch(c0["base"], c0["adj"], c0["ep"], c0["queries"], 10, 128)
    _assert_same(gpu, ora, check_ndc=False, what=f"slots={slots}")
    assert (gpu[2]["ndc"] >= ora[2]).all()
    if slots == 256:
        assert gpu[2]["lossy"].sum() > 0


@pytest.mark.parametrize("dim", [200, 96, 40, 20, 8])
@pytest.mark.parametrize("metric", [0, 1])
def test_dims_and_metrics(dim, metric):
    import oracle
    from scattersearch_amd import data
    base, q = data.mixture(2000, 40, dim, 6, 0.3, seed=dim + 100 * metric,
                           lognormal_sigma=0.5 if metric else 0.0, unit_norm=bool(metric))
    adj, ep, _ = oracle.cpu_graph(base, metric, 20, 20, 24, passes=1)
    ix = _index(base, adj, ep, metric)
    gpu = ix.search(q, 10, 40, visited_slots=8192)
    ora = oracle.search_batch(base, adj, ep, q, 10, 40, metric=metric)
    _assert_same(gpu, ora, what=f"dim={dim} metric={metric}")


@pytest.mark.parametrize("seed", range(6))
def test_random_graphs_with_ties(seed):
    """Random (non-navigable) graphs over data with many repeated vectors, so
    that exact distance ties reach the pool order, the tail filter and the
    merge."""
    import oracle
    rng = np.random.default_rng(seed)
    num = 600
    distinct = rng.standard_normal((60, 16)).astype(np.float32)
    base = distinct[rng.integers(0, 60, size=num)]
    if seed % 2:
        base = np.round(base * 2) / 2  # coarse grid: more ties
    q = rng.standard_normal((30, 16)).astype(np.float32)
    r_pad = int(rng.choice([8, 32, 64]))
    adj = random_graph(rng, num, r_pad)
    ep = int(rng.integers(0, num))
    ix = _index(base, adj, ep)
    for L in (4, 16, 70):
        K = min(L, 8)
        gpu = ix.search(q, K, L, visited_slots=4096)
        ora = oracle.search_batch(base, adj, ep, q, K, L)
        _assert_same(gpu, ora, what=f"seed={seed} L={L}")


@pytest.mark.parametrize("K,L", [(1, 1), (1, 2), (2, 2), (64, 64), (65, 65), (512, 512), (512, 1024)])
def test_extreme_k_and_l_equal_oracle(c0, K, L):
    """The smallest pools, K = L at chunk boundaries, and K at its maximum
    (SS_MAX_K = 512) for the register, LDS and large-pool kernels."""
    import oracle
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    q = c0["queries"][:24]
    print(f"K={K} L={L}", flush=True)
    ids, dists, st = ix.search(q, K, L, visited_slots=8192)
    o_ids, o_d, o_ndc, o_hops = oracle.search_batch(c0["base"], c0["adj"], c0["ep"], q, K, L)
    assert np.array_equal(ids, o_ids), f"K={K} L={L}: ids differ"
    assert np.array_equal(dists.view(np.uint32), o_d.view(np.uint32)), f"K={K} L={L}: distances differ"
    assert np.array_equal(st["hops"], o_hops)
    exact = st["lossy"] == 0
    assert np.array_equal(st["ndc"][exact], o_ndc[exact])
    assert (st["status"] == 0).all()


def test_small_entry_degree_fills_with_none():
    """deg(ep) < K: only improving candidates grow the pool, unfilled slots
    come back as (SS_NONE, +inf)."""
    import oracle
    rng = np.random.default_rng(11)
    base = rng.standard_normal((50, 16)).astype(np.float32)
    adj = np.full((50, 4), NONE, dtype=np.uint32)
    adj[0, :2] = [1, 2]  # entry point with two neighbours, nodes 1, 2 lead nowhere
    ix = _index(base, adj, 0)
    ids, d, st = ix.search(base[5:8], 10, 20)
    o = oracle.search_batch(base, adj, 0, base[5:8], 10, 20)
    assert np.array_equal(ids, o[0])
    assert (ids[:, 2:] == NONE).all() and np.isinf(d[:, 2:]).all()


def _components_graph(rng, sizes, r_pad):
    """Disconnected graph: nodes split into components of the given sizes,
    each a random graph over its own nodes with degrees from 0 up, some rows
    holding their own id (a self-loop).  Returns (adj, list of node arrays)."""
    num = int(sum(sizes))
    perm = rng.permutation(num).astype(np.uint32)
    adj = np.full((num, r_pad), NONE, dtype=np.uint32)
    comps, at = [], 0
    for size in sizes:
        nodes = perm[at:at + size]
        at += size
        comps.append(nodes)
        for n in nodes:
            d = int(rng.integers(0, min(r_pad, size) + 1))
            row = rng.choice(nodes, size=d, replace=False)
            if d and rng.random() < 0.2 and n not in row:
                row[0] = n
            adj[n, :d] = row
    return adj, comps


@pytest.mark.parametrize("seed", range(3))
def test_disconnected_components_equal_oracle(seed):
    """Entry points in components of 1 to 700 nodes, rows of degree 0 and
    self-loops: a search ends once the reachable nodes are all expanded,
    with fewer than K (or L) of them in the pool; both pool kernels, teams of
    4 and 8, every pool layout and visited-table shape give the oracle's
    ids, distance bits, distance and hop counts, SS_NONE past the filled
    pool (src/annsearch.cpp:182-259 ends the same way)."""
    import oracle
    from scattersearch_amd._lib import (SS_POOL_LDS, SS_POOL_REGISTERS, SS_POOL_SHIFT,
                                        SS_VISITED_TWO_CHOICE)
    rng = np.random.default_rng(100 + seed)
    r_pad = int(rng.choice([8, 16, 32]))
    sizes = [1, 1, 3, 9, 40, 300, 700]
    adj, comps = _components_graph(rng, sizes, r_pad)
    num = adj.shape[0]
    dim = int(rng.choice([16, 100]))
    base = rng.standard_normal((num, dim)).astype(np.float32)
    q = rng.standard_normal((24, dim)).astype(np.float32)
    if seed == 0:  # one isolated node without its self-loop: an empty first pool
        adj[comps[0][0]] = NONE
    shapes = [dict(), dict(team_size=4), dict(team_size=8),
              dict(pool_layout=SS_POOL_REGISTERS), dict(pool_layout=SS_POOL_LDS),
              dict(pool_layout=SS_POOL_SHIFT, team_size=4, waves_per_block=1),
              dict(visited_slots=2048, visited_bucket=16 | SS_VISITED_TWO_CHOICE)]
    for c, nodes in enumerate(comps):
        ep = int(nodes[rng.integers(0, nodes.size)])
        ix = _index(base, adj, ep)
        for L in (4, 16, 70, 600):
            K = min(L, 10)
            ora = oracle.search_batch(base, adj, ep, q, K, L)
            for shape in shapes:
                if shape.get("pool_layout") == SS_POOL_REGISTERS and L > 64:
                    continue
                what = f"seed={seed} component of {nodes.size} L={L} {shape}"
                print(what, flush=True)
                gpu = ix.search(q, K, L, **shape)
                exact = (gpu[2]["lossy"] == 0).all()
                _assert_same(gpu, ora, check_ndc=exact, what=what)
                filled = (ora[0] != NONE).sum(1)
                assert (filled <= nodes.size).all()
                assert np.isinf(gpu[1][ora[0] == NONE]).all(), what


def test_hipgraph_replay_matches_eager(small):
    ix = _index(small["base"], small["adj"], small["ep"])
    eager = ix.search(small["queries"], 10, 48)
    for _ in range(4):
        g = ix.search(small["queries"], 10, 48, use_graph=True)
        assert np.array_equal(eager[0], g[0])
        assert np.array_equal(eager[2]["hops"], g[2]["hops"])
        mid = ix.search(small["queries"][:7], 5, 20)  # eager work between replays
        assert mid[0].shape == (7, 5)


def test_device_buffers_api(small):
    torch = pytest.importorskip("torch")
    from scattersearch_amd import dim_padded, STATS_DTYPE
    ix = _index(small["base"], small["adj"], small["ep"])
    q = small["queries"]
    nq, dim = q.shape
    dq = torch.zeros((nq, dim_padded(dim)), dtype=torch.float32, device="cuda")
    dq[:, :dim] = torch.from_numpy(q).cuda()
    ids = torch.empty((nq, 16), dtype=torch.int32, device="cuda")
    dd = torch.empty((nq, 16), dtype=torch.float32, device="cuda")
    st = torch.empty((nq, 8), dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream()
    ix.search_device(dq.data_ptr(), nq, 10, 48, 16, ids.data_ptr(), dd.data_ptr(), st.data_ptr(),
                     stream=s.cuda_stream)
    torch.cuda.synchronize()
    host = ix.search(q, 10, 48, k_stride=16)
    assert np.array_equal(ids.cpu().numpy().view(np.uint32), host[0])
    dev_stats = st.cpu().numpy().view(np.uint32).reshape(-1).view(STATS_DTYPE)
    for f in ("ndc", "hops", "lossy"):
        assert np.array_equal(dev_stats[f], host[2][f])


def test_invalid_arguments_rejected(small):
    from scattersearch_amd import SSError
    ix = _index(small["base"], small["adj"], small["ep"])
    with pytest.raises(SSError):
        ix.search(small["queries"], 20, 10)  # K > L
    with pytest.raises(SSError):
        ix.search(small["queries"], 10, 16385)  # L above the pool bound
    with pytest.raises(SSError):
        ix.search(small["queries"], 10, 48, visited_slots=1000)  # not a power of two
    bad = small["adj"].copy()
    bad[3, 0] = small["base"].shape[0] + 5
    with pytest.raises(SSError):
        ix.set_graph(bad, small["ep"])


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("seed", range(4))
def test_wide_expansion_matches_model(width, seed):
    """Expanding the first 2 or 4 unexplored entries per step: the device
    result equals the step model of tests/merge_model.py exactly."""
    import oracle
    from merge_model import merge_search
    rng = np.random.default_rng(50 + seed)
    num = 300
    base = rng.integers(-3, 4, (num, 16)).astype(np.float32) if seed % 2 else \
        rng.standard_normal((num, 16)).astype(np.float32)
    adj = random_graph(rng, num, 12)
    ep = int(rng.integers(0, num))
    q = rng.standard_normal((12, 16)).astype(np.float32)
    ix = _index(base, adj, ep)
    for L in (8, 40):
        ids, d, st = ix.search(q, 8, L, visited_slots=4096, expand_width=width)
        for i in range(q.shape[0]):
            m = merge_search(base, adj, ep, q[i], 8, L, lambda x, y: np.float32(oracle.dist(0, x, y)),
                             width=width)
            assert np.array_equal(ids[i], m[0]), (seed, L, i)
            assert np.array_equal(d[i].view(np.uint32), m[1].view(np.uint32))
            assert st["ndc"][i] == m[2] and st["hops"][i] == m[3]


@pytest.mark.parametrize("width", [2, 4])
def test_wide_expansion_recall(c0, width):
    import oracle
    from scattersearch_amd import metrics
    gt, _ = oracle.bruteforce(c0["base"], c0["queries"], 10)
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    exact = ix.search(c0["queries"], 10, 64)
    wide = ix.search(c0["queries"], 10, 64, expand_width=width)
    r1, rw = metrics.recall(exact[0], gt, 10), metrics.recall(wide[0], gt, 10)
    assert rw >= r1 - 0.02
    assert wide[2]["hops"].mean() >= exact[2]["hops"].mean() * 0.8  # same work, fewer steps
    for r in range(wide[0].shape[0]):
        assert (np.diff(wide[1][r]) >= 0).all() and len(set(wide[0][r].tolist())) == 10


def test_autotune_exact_and_relaxed(small):
    import oracle
    from scattersearch_amd import metrics
    ix = _index(small["base"], small["adj"], small["ep"])
    exact = ix.search(small["queries"], 10, 48, expand_width=1)
    best = ix.autotune(small["queries"], 10, 48, max_recall_loss=0.0)
    assert best["expand_width"] == 1 and best["ms"] > 0
    # a search of another form does not take the tuned shape's team size
    # (teams of 8 only run one node per step on one wavefront)
    for kw in (dict(expand_width=2), dict(waves_per_query=2)):
        ids = ix.search(small["queries"], 10, 48, **kw)[0]
        assert (ids[:, 0] != NONE).all(), kw
    tuned = ix.search(small["queries"], 10, 48)  # parameters left at 0 take the tuned ones
    assert np.array_equal(tuned[0], exact[0])
    gt, _ = oracle.bruteforce(small["base"], small["queries"], 10)
    r_exact = metrics.recall(exact[0], gt, 10)
    assert abs(best["recall"] - r_exact) < 1e-6
    relaxed = ix.autotune(small["queries"], 10, 48, max_recall_loss=0.05)
    assert relaxed["recall"] >= r_exact - 0.05 - 1e-6
    assert relaxed["expand_width"] in (1, 2, 4)
    assert relaxed["waves_per_query"] in (1, 2, 4)
    r_tuned = metrics.recall(ix.search(small["queries"], 10, 48)[0], gt, 10)  # relaxed parameters apply
    assert r_tuned >= r_exact - 0.05 - 1e-6


def test_small_16bit_table_with_lossy_inserts():
    """2048-d vectors and L=512 leave room for fewer than 16 wavefronts per CU,
    so the launch keeps 16-bit visited slots; 256 slots make lanes of one step
    collide on slots and fill buckets (lossy inserts).  Ids, distance bits and
    hops still equal the oracle's; distance counts wherever nothing was lost."""
    import oracle
    from scattersearch_amd import data
    base, q = data.mixture(3000, 40, 2048, 8, 0.3, seed=2048)
    adj, ep, _ = oracle.cpu_graph(base, 0, 24, 24, 32, passes=1)
    ix = _index(base, adj, ep)
    gpu = ix.search(q, 10, 512, visited_slots=256)
    assert gpu[2]["lossy"].sum() > 0
    ora = oracle.search_batch(base, adj, ep, q, 10, 512)
    _assert_same(gpu, ora, check_ndc=False, what="16-bit table, 256 slots")
    exact = gpu[2]["lossy"] == 0
    assert np.array_equal(gpu[2]["ndc"][exact], ora[2][exact])
    assert (gpu[2]["ndc"] >= ora[2]).all()


@pytest.mark.parametrize("slots", [256, 1024, 4096])
@pytest.mark.parametrize("L", [48, 128, 512])
def test_wide_16bit_buckets_equal_oracle(c0, slots, L):
    """The 16-bit visited table in buckets of sixteen slots (two 16-byte reads
    per lookup): ids, distance bits and hops equal the oracle at every pool
    layout; 256 slots fill buckets (lossy inserts, recomputed distances
    only), 4,096 forget nothing up to L = 128, so distance counts match."""
    import oracle
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    q = c0["queries"]
    gpu = ix.search(q, 10, L, visited_slots=slots, visited_bucket=16)
    ora = oracle.search_batch(c0["base"], c0["adj"], c0["ep"], q, 10, L)
    _assert_same(gpu, ora, check_ndc=False, what=f"16-slot buckets, {slots} slots, L={L}")
    st = gpu[2]
    exact = st["lossy"] == 0
    assert np.array_equal(st["ndc"][exact], ora[2][exact])
    assert (st["ndc"] >= ora[2]).all()
    if slots == 256 and L >= 128:
        assert st["lossy"].sum() > 0
    if slots == 4096 and L <= 128:
        assert st["lossy"].sum() == 0


def test_wide_buckets_forget_less_than_narrow(c0):
    """At the same table size, buckets of sixteen slots overflow far less
    often than buckets of eight."""
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    q = c0["queries"]
    narrow = ix.search(q, 10, 256, visited_slots=1024, visited_bucket=8)[2]["lossy"].sum()
    wide = ix.search(q, 10, 256, visited_slots=1024, visited_bucket=16)[2]["lossy"].sum()
    assert wide < narrow, (wide, narrow)


@pytest.mark.parametrize("bucket", [8, 16])
@pytest.mark.parametrize("slots", [256, 1024, 4096])
@pytest.mark.parametrize("L", [48, 128, 512, 900])
def test_two_choice_buckets_equal_oracle(c0, bucket, slots, L):
    """The 16-bit table with a second bucket per id (SS_VISITED_TWO_CHOICE):
    ids, distance bits and hops equal the oracle for both bucket sizes, at
    every pool layout and in the large-pool kernel (L = 900); distance
    counts equal it wherever nothing was forgotten and never fall below."""
    import oracle
    from scattersearch_amd._lib import SS_VISITED_TWO_CHOICE
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    q = c0["queries"]
    gpu = ix.search(q, 10, L, visited_slots=slots, visited_bucket=bucket | SS_VISITED_TWO_CHOICE)
    ora = oracle.search_batch(c0["base"], c0["adj"], c0["ep"], q, 10, L)
    _assert_same(gpu, ora, check_ndc=False, what=f"two-choice buckets of {bucket}, {slots} slots, L={L}")
    st = gpu[2]
    exact = st["lossy"] == 0
    assert np.array_equal(st["ndc"][exact], ora[2][exact])
    assert (st["ndc"] >= ora[2]).all()
    if slots == 256 and L >= 128:
        assert st["lossy"].sum() > 0


def test_two_choice_buckets_forget_less_than_one(c0):
    """At the same table size and load, a second bucket per id forgets far
    fewer ids than one bucket."""
    from scattersearch_amd._lib import SS_VISITED_TWO_CHOICE
    ix = _index(c0["base"], c0["adj"], c0["ep"])
    q = c0["queries"]
    for bucket in (8, 16):
        one = ix.search(q, 10, 256, visited_slots=1024, visited_bucket=bucket)[2]["lossy"].sum()
        two = ix.search(q, 10, 256, visited_slots=1024, visited_bucket=bucket | SS_VISITED_TWO_CHOICE)[2]
        assert two["lossy"].sum() < one, (bucket, two["lossy"].sum(), one)


def test_two_choice_buckets_in_one_launch_mixing_pool_sizes(small):
    """Per-query K and L through the pool-size classes and through one
    kernel with two-choice buckets equal the oracle."""
    import oracle
    from scattersearch_amd._lib import SS_CLASSES_AUTO, SS_CLASSES_OFF, SS_VISITED_TWO_CHOICE
    rng = np.random.default_rng(29)
    q = small["queries"]
    nq = q.shape[0]
    K = rng.choice([1, 10, 32], size=nq).astype(np.uint32)
    L = np.array([rng.integers(max(int(k), 16), 513) for k in K], dtype=np.uint32)
    ix = _index(small["base"], small["adj"], small["ep"])
    ora = oracle.search_batch(small["base"], small["adj"], small["ep"], q, K, L, k_stride=32)
    for mode in (SS_CLASSES_AUTO, SS_CLASSES_OFF):
        gpu = ix.search(q, K, L, k_stride=32, visited_bucket=16 | SS_VISITED_TWO_CHOICE, visited_slots=2048,
                        pool_classes=mode)
        for i in range(nq):
            k = int(K[i])
            assert np.array_equal(gpu[0][i, :k], ora[0][i, :k]), f"mode {mode} query {i} K={k} L={L[i]}"
            assert np.array_equal(gpu[1][i, :k].view(np.uint32), ora[1][i, :k].view(np.uint32))
            assert gpu[2]["hops"][i] == ora[3][i]


def test_wide_buckets_in_one_launch_mixing_pool_sizes(small):
    """Per-query K and L through the pool-size classes with 16-slot buckets
    equal the oracle."""
    import oracle
    rng = np.random.default_rng(23)
    q = small["queries"]
    nq = q.shape[0]
    K = rng.choice([1, 10, 32], size=nq).astype(np.uint32)
    L = np.array([rng.integers(max(int(k), 16), 513) for k in K], dtype=np.uint32)
    ix = _index(small["base"], small["adj"], small["ep"])
    gpu = ix.search(q, K, L, k_stride=32, visited_bucket=16, visited_slots=2048)
    ora = oracle.search_batch(small["base"], small["adj"], small["ep"], q, K, L, k_stride=32)
    for i in range(nq):
        k = int(K[i])
        assert np.array_equal(gpu[0][i, :k], ora[0][i, :k]), f"query {i} K={k} L={L[i]}"
        assert np.array_equal(gpu[1][i, :k].view(np.uint32), ora[1][i, :k].view(np.uint32))
        assert gpu[2]["hops"][i] == ora[3][i]


def test_visited_bucket_rejects_other_sizes(small):
    from scattersearch_amd import SSError
    from scattersearch_amd._lib import SS_EINVAL
    from scattersearch_amd._lib import SS_VISITED_TWO_CHOICE
    ix = _index(small["base"], small["adj"], small["ep"])
    for bad in (4, 32, SS_VISITED_TWO_CHOICE, 4 | SS_VISITED_TWO_CHOICE, 16 | 0x200):
        with pytest.raises(SSError) as e:
            ix.search(small["queries"][:4], 10, 48, visited_bucket=bad)
        assert e.value.code == SS_EINVAL, bad
