"""Pure-Python restatement of the device search's expand-then-merge step, for
tiny instances.  Used by the CPU tests to check that the whole-row merge
reproduces the sequential reference loop (src/annsearch.cpp:218-253 with
InsertIntoPool, include/util.h:41-78) exactly, ties included, and that a
forgetful visited set changes nothing but the distance count.
"""
import numpy as np

NONE = 0xFFFFFFFF


def merge_search(base, adj, ep, q, K, L, dist, forget=None, width=1):
    """dist(a, b) -> float32 distance.  forget(id) -> True to treat a visited
    id as unvisited (a lossy visited set).  width > 1 expands the first
    `width` unexplored entries per step against the tail at the start of the
    step (the device's wider expansion).  Returns ids, dists, ndc, hops."""
    if width > 1:
        return _wide_search(base, adj, ep, q, K, L, dist, width)
    row = [int(x) for x in adj[ep] if x != NONE]
    t = min(L, len(row))
    visited = set(row[:t])
    seeds = sorted(((dist(base[i], q), j, i) for j, i in enumerate(row[:t])),
                   key=lambda s: (s[0], s[1]))
    pool = [[d, i, False] for d, _, i in seeds]
    ndc, hops, start = t, 0, 0
    while True:
        k = next((i for i in range(start, len(pool)) if not pool[i][2]), None)
        if k is None:
            break
        pool[k][2] = True
        n = pool[k][1]
        hops += 1
        tail = pool[-1][0]
        A = []
        for m, x in enumerate(adj[n]):
            x = int(x)
            if x == NONE:
                break
            if x in visited and not (forget and forget(x)):
                continue
            visited.add(x)
            d = dist(base[x], q)
            ndc += 1
            if d < tail and not any(p[1] == x for p in pool):
                A.append((d, m, x))
        if not A:
            start = k + 1
            continue
        A.sort(key=lambda a: (a[0], a[1]))
        merged, i, j = [], 0, 0
        pos_first = None
        while i < len(pool) or j < len(A):
            if j == len(A) or (i < len(pool) and pool[i][0] <= A[j][0]):
                merged.append(pool[i])
                i += 1
            else:
                if j == 0:
                    pos_first = len(merged)
                merged.append([A[j][0], A[j][2], False])
                j += 1
        pool = merged[:L]
        start = pos_first if pos_first <= k else k + 1
    ids = np.full(K, NONE, dtype=np.uint32)
    ds = np.full(K, np.inf, dtype=np.float32)
    for i in range(min(K, len(pool))):
        ids[i] = pool[i][1]
        ds[i] = pool[i][0]
    return ids, ds, ndc, hops


def _wide_search(base, adj, ep, q, K, L, dist, width):
    row = [int(x) for x in adj[ep] if x != NONE]
    t = min(L, len(row))
    visited = set(row[:t])
    seeds = sorted(((dist(base[i], q), j, i) for j, i in enumerate(row[:t])),
                   key=lambda s: (s[0], s[1]))
    pool = [[d, i, False] for d, _, i in seeds]
    ndc, hops = t, 0
    while True:
        ks = [i for i in range(len(pool)) if not pool[i][2]][:width]
        if not ks:
            break
        tail = pool[-1][0]
        A = []
        order = 0
        for k in ks:
            pool[k][2] = True
            hops += 1
        for k in ks:
            for x in adj[pool[k][1]]:
                x = int(x)
                if x == NONE:
                    break
                if x in visited:
                    continue
                visited.add(x)
                d = dist(base[x], q)
                ndc += 1
                if d < tail and not any(p[1] == x for p in pool):
                    A.append((d, order, x))
                order += 1
        if not A:
            continue
        A.sort(key=lambda a: (a[0], a[1]))
        merged, i, j = [], 0, 0
        while i < len(pool) or j < len(A):
            if j == len(A) or (i < len(pool) and pool[i][0] <= A[j][0]):
                merged.append(pool[i])
                i += 1
            else:
                merged.append([A[j][0], A[j][2], False])
                j += 1
        pool = merged[:L]
    ids = np.full(K, NONE, dtype=np.uint32)
    ds = np.full(K, np.inf, dtype=np.float32)
    for i in range(min(K, len(pool))):
        ids[i] = pool[i][1]
        ds[i] = pool[i][0]
    return ids, ds, ndc, hops
