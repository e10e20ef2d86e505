"""Result metrics with the reference drivers' definitions.

recall@K: |res[0..K) ∩ gt[0..K)| / K per query, averaged
(tests/search.cpp:91-107).  p-th percentile of a list: the ascending-sorted
list at index floor(n * p) (tests/parallel_search.cpp:94-95,
tests/search.cpp:108-110).
"""
import numpy as np


def recall_per_query(res, gt, K):
    res = np.asarray(res)[:, :K]
    gt = np.asarray(gt)[:, :K]
    hits = np.zeros(res.shape[0], dtype=np.int64)
    for j in range(K):
        hits += (res[:, j:j + 1] == gt).any(axis=1)
    return hits / float(K)


def recall(res, gt, K):
    return float(recall_per_query(res, gt, K).mean())


def percentile(values, p):
    v = np.sort(np.asarray(values, dtype=np.float64))
    if v.size == 0:
        return float("nan")
    return float(v[min(int(v.size * p), v.size - 1)])


def bytes_per_query(ndc, hops, dim, r_pad, K):
    """Algorithmic HBM bytes of one query: gathered base vectors, adjacency
    rows of expanded nodes, the query vector and the K (id, distance) written
    back."""
    return (np.asarray(ndc, dtype=np.float64) * dim * 4 + np.asarray(hops, dtype=np.float64) * r_pad * 4
            + dim * 4 + np.asarray(K, dtype=np.float64) * 8)
