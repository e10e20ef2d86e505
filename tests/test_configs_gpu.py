"""Parity cases shaped like the other workloads of the benchmark definition,
at sizes the CPU oracle checks in seconds:

* inner product on unit-normalised 200-d data with log-normal norms, R=64,
  and one launch mixing K in {1, 10, 100} with L uniform in K..512;
* 96-d data, R=32, persistent hipGraph-replayed launches of batch sizes
  1, 64 and 1024.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ip200():
    import oracle
    from scattersearch_amd import data
    base, q = data.mixture(12000, 300, 200, 40, 0.25, seed=77, lognormal_sigma=0.5, unit_norm=True)
    adj, ep, _ = oracle.cpu_graph(base, 1, 64, 64, 64, passes=2)
    return dict(base=base, queries=q, adj=adj, ep=ep)


@pytest.fixture(scope="module")
def d96():
    import oracle
    from scattersearch_amd import data
    base, q = data.mixture(15000, 1024, 96, 60, 0.2, seed=96)
    adj, ep, _ = oracle.cpu_graph(base, 0, 32, 32, 64, passes=2)
    return dict(base=base, queries=q, adj=adj, ep=ep)


def test_ip_d200_r64_mixed_k_and_l_one_launch(ip200):
    import oracle
    from scattersearch_amd import Index, metrics
    d = ip200
    rng = np.random.default_rng(2)
    nq = d["queries"].shape[0]
    K = np.array([1, 10, 100] * (nq // 3), dtype=np.uint32)
    L = np.array([rng.integers(k, 513) for k in K], dtype=np.uint32)
    ix = Index(d["base"], 1)
    ix.set_graph(d["adj"], d["ep"])
    assert ix.info()["r_pad"] == 64
    ids, ds, st = ix.search(d["queries"], K, L, k_stride=100, visited_slots=16384)
    o_ids, o_ds, o_ndc, o_hops = oracle.search_batch(d["base"], d["adj"], d["ep"], d["queries"], K, L,
                                                     metric=1, k_stride=100)
    assert np.array_equal(ids, o_ids)
    assert np.array_equal(ds.view(np.uint32), o_ds.view(np.uint32))
    assert np.array_equal(st["hops"], o_hops)
    exact = st["lossy"] == 0
    assert np.array_equal(st["ndc"][exact], o_ndc[exact])
    gt, _ = ix.bruteforce(d["queries"], 100)
    for k in (10, 100):
        sel = K == k
        assert metrics.recall(ids[sel], gt[sel], k) > 0.8


@pytest.mark.parametrize("batch", [1, 64, 1024])
def test_d96_persistent_graph_batches(d96, batch):
    import oracle
    from scattersearch_amd import Index
    d = d96
    ix = Index(d["base"])
    ix.set_graph(d["adj"], d["ep"])
    q = d["queries"][:batch]
    o = oracle.search_batch(d["base"], d["adj"], d["ep"], q, 10, 64)
    for rep in range(3):  # replays of the captured launch
        ids, ds, st = ix.search(q, 10, 64, use_graph=True, visited_slots=8192)
        assert np.array_equal(ids, o[0])
        assert np.array_equal(ds.view(np.uint32), o[1].view(np.uint32))
        assert np.array_equal(st["hops"], o[3])


def test_ip_d200_scatter_recall(ip200):
    """Several wavefronts per query on the inner-product workload, judged by
    recall@10 against exact ground truth."""
    from scattersearch_amd import Index, metrics
    d = ip200
    ix = Index(d["base"], 1)
    ix.set_graph(d["adj"], d["ep"])
    gt, _ = ix.bruteforce(d["queries"], 10)
    exact = ix.search(d["queries"], 10, 128, visited_slots=16384)[0]
    ids, ds, st = ix.search(d["queries"], 10, 128, waves_per_query=4, visited_slots=16384)
    assert metrics.recall(ids, gt, 10) >= metrics.recall(exact, gt, 10) - 0.03
    assert (np.diff(ds[:, :10], axis=1) >= 0).all()
