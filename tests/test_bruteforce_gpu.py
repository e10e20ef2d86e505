"""GPU brute force (ground truth and k-NN graph) against the CPU oracle:
integer ids and distances must be bit-identical."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("metric,dim", [(0, 128), (1, 200), (0, 96), (0, 24), (1, 40)])
@pytest.mark.parametrize("k", [1, 10, 100])
def test_bruteforce_bit_exact(metric, dim, k):
    import oracle
    from scattersearch_amd import Index, data
    base, q = data.mixture(5000, 70, dim, 10, 0.3, seed=dim * 3 + k, unit_norm=bool(metric))
    ix = Index(base, metric)
    ids, d = ix.bruteforce(q, k)
    o_ids, o_d = oracle.bruteforce(base, q, k, metric)
    assert np.array_equal(ids, o_ids)
    assert np.array_equal(d.view(np.uint32), o_d.view(np.uint32))


def test_bruteforce_ties_break_by_id():
    import oracle
    from scattersearch_amd import Index
    rng = np.random.default_rng(5)
    distinct = rng.standard_normal((30, 32)).astype(np.float32)
    base = distinct[rng.integers(0, 30, size=3000)]  # every vector repeated ~100x
    q = rng.standard_normal((20, 32)).astype(np.float32)
    ix = Index(base)
    ids, d = ix.bruteforce(q, 128)
    o_ids, o_d = oracle.bruteforce(base, q, 128)
    assert np.array_equal(ids, o_ids)
    # ids inside each run of equal distances ascend
    for r in range(ids.shape[0]):
        same = d[r, 1:] == d[r, :-1]
        assert (ids[r, 1:][same] > ids[r, :-1][same]).all()


def test_knn_graph_excludes_self():
    import oracle
    from scattersearch_amd import Index, data
    base, _ = data.mixture(4000, 1, 128, 12, 0.3, seed=9)
    ix = Index(base)
    ids, d = ix.knn_graph(32)
    o_ids, o_d = oracle.bruteforce(base, base, 32, exclude=np.arange(4000, dtype=np.uint32))
    assert np.array_equal(ids, o_ids)
    assert np.array_equal(d.view(np.uint32), o_d.view(np.uint32))
    assert not (ids == np.arange(4000)[:, None]).any()


def test_bruteforce_fewer_points_than_k():
    from scattersearch_amd import Index
    rng = np.random.default_rng(1)
    base = rng.standard_normal((5, 16)).astype(np.float32)
    ix = Index(base)
    ids, d = ix.bruteforce(base[:2], 8)
    assert (ids[:, 5:] == 0xFFFFFFFF).all() and np.isinf(d[:, 5:]).all()
    assert sorted(ids[0, :5].tolist()) == list(range(5))


def test_nn_descent_lists(small):
    """NN-descent k-NN lists: sorted, no self, no repeats, every distance the
    exact one, and close to the exact k-NN lists."""
    import oracle
    from scattersearch_amd import Index
    base = small["base"]
    ix = Index(base)
    k = 16
    ids, ds, upd = ix.nn_descent(k, iters=8, seed=5)
    assert upd.shape == (8,) and upd[0] > 0
    exact, _ = ix.knn_graph(k)
    hit = 0
    for u in range(base.shape[0]):
        row = ids[u]
        assert u not in row and len(set(row.tolist())) == k
        assert (np.diff(ds[u]) >= 0).all()
        for j in (0, k - 1):
            assert np.float32(oracle.dist(0, base[row[j]], base[u])).view(np.uint32) == ds[u, j].view(np.uint32)
        hit += len(set(row.tolist()) & set(exact[u].tolist()))
    assert hit / (base.shape[0] * k) > 0.9
