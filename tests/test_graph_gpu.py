"""Graph construction with the occlusion prunes on the GPU
(ss_build_graph_gpu): the graph must be identical to the CPU builder's
(ss_build_graph) on the same candidates - both metrics, repeated, invalid and
self ids in the lists, the relaxed rule (alpha > 1), and a hub whose reverse
list is too long for the device's lists (pruned on the host).  Then the
10M x 96 workload's pipeline at 1M points: NN-descent k-NN lists, navigation
passes and GPU prunes, and the search on that graph against the oracle.
(The reference reads graphs built by the external NSG tool,
src/annsearch.cpp:64-80; the search it runs is src/annsearch.cpp:182-259.)"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NONE = 0xFFFFFFFF


def _both(base, metric, knn, extra, R, ep, alpha=1.0):
    from scattersearch_amd import Index
    from scattersearch_amd.graph import prune_and_connect, prune_and_connect_gpu
    ix = Index(base, metric)
    cpu = prune_and_connect(base, metric, knn, extra, R, ep, alpha)
    gpu = prune_and_connect_gpu(ix, base, knn, extra, R, ep, alpha)
    return cpu, gpu


@pytest.mark.parametrize("metric", [0, 1])
def test_gpu_prune_equals_cpu_builder(metric):
    import oracle
    from scattersearch_amd import data
    from scattersearch_amd.graph import medoid
    dim = 128 if metric == 0 else 200
    base, _ = data.mixture(3000, 1, dim, 16, 0.3, seed=41 + metric, lognormal_sigma=0.5 * metric,
                           unit_norm=bool(metric))
    knn = oracle.bruteforce(base, base, 24, metric, exclude=np.arange(3000, dtype=np.uint32))[0]
    ep = medoid(base)
    r = oracle.search_batch(base, knn, ep, base, 32, 32, metric, path_cap=64)
    extra = np.concatenate([r[-1], r[0]], axis=1)
    for R, alpha in ((24, 1.0), (16, 1.2)):
        (a_cpu, f_cpu), (a_gpu, f_gpu) = _both(base, metric, knn, extra, R, ep, alpha)
        assert np.array_equal(a_cpu, a_gpu), f"metric {metric} R={R} alpha={alpha}: rows differ"
        assert f_cpu == f_gpu


def test_gpu_prune_with_repeats_invalid_and_self_ids():
    rng = np.random.default_rng(5)
    num, dim = 2500, 40
    base = rng.standard_normal((num, dim)).astype(np.float32)
    knn = rng.integers(0, num, size=(num, 20)).astype(np.uint32)
    knn[:, 3] = knn[:, 1]  # repeats
    knn[rng.random((num, 20)) < 0.1] = NONE
    knn[np.arange(0, num, 7), 5] = np.arange(0, num, 7)  # the node itself
    knn[np.arange(0, num, 11), 6] = num + 3  # out of range
    extra = rng.integers(0, num, size=(num, 90)).astype(np.uint32)
    extra[:, :20] = knn  # every k-NN entry again
    (a_cpu, f_cpu), (a_gpu, f_gpu) = _both(base, 0, knn, extra, 12, 0)
    assert np.array_equal(a_cpu, a_gpu)
    assert f_cpu == f_gpu


def test_gpu_prune_hub_with_a_long_reverse_list():
    """Every node lists node 0, which sits at the centre of the data, so
    nearly every pruned list keeps it and its reverse list (thousands of
    entries) is longer than the device's candidate lists: that node is pruned
    on the host, the others on the GPU, and the graph is the CPU builder's."""
    rng = np.random.default_rng(8)
    num, dim = 3000, 16
    base = rng.standard_normal((num, dim)).astype(np.float32)
    base[0] = 0.0
    knn = rng.integers(1, num, size=(num, 10)).astype(np.uint32)
    knn[:, 0] = 0
    knn[0, 0] = 1
    (a_cpu, f_cpu), (a_gpu, f_gpu) = _both(base, 0, knn, None, 16, 0)
    assert (a_cpu == 0).sum() > 600  # the hub is kept in many lists
    assert np.array_equal(a_cpu, a_gpu)
    assert f_cpu == f_gpu


def test_nn_descent_graph_search_equals_oracle_at_1m():
    """The 10M x 96 workload's pipeline on its first 1M points (NN-descent
    k-NN lists, three navigation passes with GPU prunes) and 1,000 of its
    queries: the GPU search equals the oracle at L = 64 and 128 (ids,
    distance bits, hops, and the distance count wherever nothing was
    forgotten), the graph is connected and the recall is that of a working
    index."""
    import time
    import oracle
    from scattersearch_amd import Index, data, metrics
    from scattersearch_amd.graph import gpu_build_graph, reachable
    w = data.WORKLOADS["c3"]
    t0 = time.time()
    base, q = data.generate(w, 1_000_000, 1000)
    ix = Index(base, w.metric)
    print(f"data {time.time() - t0:.1f}s", flush=True)
    adj, ep, info = gpu_build_graph(ix, base, w.R, w.knn_k, w.nav_L, nn_descent_iters=10,
                                    log=lambda m: print(m, flush=True))
    assert reachable(adj, ep) == base.shape[0]
    gt, _ = ix.bruteforce(q, 10)
    for L in (64, 128):
        ids, ds, st = ix.search(q, 10, L)
        o_ids, o_d, o_ndc, o_hops = oracle.search_batch(base, adj, ep, q, 10, L)
        assert np.array_equal(ids, o_ids), f"L={L}: ids differ on {np.nonzero((ids != o_ids).any(1))[0][:10]}"
        assert np.array_equal(ds.view(np.uint32), o_d.view(np.uint32)), f"L={L}"
        assert np.array_equal(st["hops"], o_hops), f"L={L}"
        clean = st["lossy"] == 0
        assert np.array_equal(st["ndc"][clean], o_ndc[clean]), f"L={L}"
        rec = metrics.recall(ids, gt, 10)
        print(f"L={L}: recall@10 {rec:.4f}, mean ndc {st['ndc'].mean():.1f}, hops {st['hops'].mean():.1f}",
              flush=True)
        assert rec > 0.8
