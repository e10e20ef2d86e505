"""Golden fixtures (tests/golden/, written by make_golden.py): the
reference's own distance and InsertIntoPool outputs, and a fixed tiny search
instance."""
import os

import numpy as np
import pytest

import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


@pytest.mark.parametrize("dim", [96, 128, 200])
def test_oracle_distances_equal_reference_outputs(dim):
    g = _load("ref_vectors.npz")
    a, b = g[f"a{dim}"], g[f"b{dim}"]
    l2 = np.array([oracle.dist(0, a[i], b[i]) for i in range(a.shape[0])], np.float32)
    ip = np.array([oracle.dist(1, a[i], b[i]) for i in range(a.shape[0])], np.float32)
    assert np.array_equal(l2.view(np.uint32), g[f"l2_{dim}"].view(np.uint32))
    assert np.array_equal(ip.view(np.uint32), g[f"ip_{dim}"].view(np.uint32))


def test_oracle_insert_equals_reference_outputs():
    g = _load("ref_vectors.npz")
    i = 0
    while f"ins{i}_meta" in g:
        L, t, t1 = (int(x) for x in g[f"ins{i}_meta"])
        pool = (oracle.Neighbor * (L + 1))()
        for j in range(L + 1):
            pool[j] = oracle.Neighbor(int(g[f"ins{i}_pool_ids"][j]), float(g[f"ins{i}_pool_d"][j]), 1)
        pos = []
        for nid, d in zip(g[f"ins{i}_ids"], g[f"ins{i}_d"]):
            if d >= pool[t - 1].distance:
                pos.append(-1)
                continue
            pos.append(oracle.insert_into_pool(pool, t, oracle.Neighbor(int(nid), float(d), 1)))
            if t < L:
                t += 1
        assert pos == g[f"ins{i}_pos"].tolist(), f"sequence {i}"
        assert t == t1
        assert [p.id for p in pool] == g[f"ins{i}_out_ids"].tolist()
        i += 1
    assert i > 0


def test_oracle_reproduces_tiny_fixture():
    g = _load("tiny_search.npz")
    ids, ds, ndc, hops = oracle.search_batch(g["base"], g["adj"], int(g["ep"]), g["queries"],
                                             int(g["K"]), int(g["L"]))
    assert np.array_equal(ids, g["ids"])
    assert np.array_equal(ds.view(np.uint32), g["dists"].view(np.uint32))
    assert np.array_equal(ndc, g["ndc"]) and np.array_equal(hops, g["hops"])
    gt, gd = oracle.bruteforce(g["base"], g["queries"], 10)
    assert np.array_equal(gt, g["gt"])


def test_tiny_fixture_recall_is_high():
    from scattersearch_amd import metrics
    g = _load("tiny_search.npz")
    assert metrics.recall(g["ids"], g["gt"], 10) > 0.9


@pytest.mark.gpu
def test_gpu_reproduces_tiny_fixture():
    from scattersearch_amd import Index
    g = _load("tiny_search.npz")
    ix = Index(g["base"])
    ix.set_graph(g["adj"], int(g["ep"]))
    ids, ds, st = ix.search(g["queries"], int(g["K"]), int(g["L"]), visited_slots=8192)
    assert np.array_equal(ids, g["ids"])
    assert np.array_equal(ds.view(np.uint32), g["dists"].view(np.uint32))
    assert np.array_equal(st["ndc"], g["ndc"]) and np.array_equal(st["hops"], g["hops"])
    gt, gd = ix.bruteforce(g["queries"], 10)
    assert np.array_equal(gt, g["gt"]) and np.array_equal(gd.view(np.uint32), g["gt_dists"].view(np.uint32))
