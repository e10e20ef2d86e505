"""NSG-style graph construction for the synthetic workloads.

The reference only reads .nsg files made by the external NSG tool
(src/annsearch.cpp:64-80).  This pipeline makes them:

1. exact k-NN of every base point (GPU brute force, bit-exact with the CPU);
2. entry point = the point nearest the centroid (NSG's navigating node);
3. navigation candidates: the pool of a search for every point from the
   entry point over the k-NN graph (NSG takes candidates from this search
   path, which is what gives the graph its long-range edges between
   clusters);
4. occlusion prune of k-NN + navigation candidates to R (MRNG with the
   relaxation factor alpha of DiskANN's RobustPrune), reverse edges, and a
   pass that attaches every node not reachable from the entry point
   (ss_build_graph in the C ABI, or ss_build_graph_gpu with both prunes on
   the GPU and the same result);
5. steps 3-4 again with the navigation search over the graph of step 4.
"""
import ctypes

import numpy as np

from ._lib import check, fptr, lib, uptr


def medoid(base):
    base = np.ascontiguousarray(base, dtype=np.float32)
    ep = ctypes.c_uint32()
    check(lib.ss_medoid(fptr(base), base.shape[0], base.shape[1], ctypes.byref(ep)))
    return ep.value


def prune_and_connect(base, metric, knn, extra, R, ep, alpha=1.0, nthreads=0):
    base = np.ascontiguousarray(base, dtype=np.float32)
    knn = np.ascontiguousarray(knn, dtype=np.uint32)
    ex = None if extra is None else np.ascontiguousarray(extra, dtype=np.uint32)
    adj = np.empty((base.shape[0], R), dtype=np.uint32)
    fixed = ctypes.c_uint32()
    check(lib.ss_build_graph(fptr(base), base.shape[0], base.shape[1], int(metric), uptr(knn),
                             knn.shape[1], uptr(ex), 0 if ex is None else ex.shape[1], R, alpha, ep,
                             nthreads, uptr(adj), ctypes.byref(fixed)))
    return adj, fixed.value


def reachable(adj, ep):
    a = np.ascontiguousarray(adj, dtype=np.uint32)
    r = ctypes.c_uint32()
    check(lib.ss_graph_reachable(uptr(a), a.shape[0], a.shape[1], ep, ctypes.byref(r)))
    return r.value


def build_graph(base, metric, R, knn_k, nav_L, knn_fn, search_fn, alpha=1.0, passes=3,
                nthreads=0, log=None, prune_fn=None):
    """knn_fn(k) -> [num][k] ids; search_fn(adj, ep, L, path_cap, out_pool,
    out_path) fills out_pool [num][L] with the pool ids and out_path
    [num][path_cap] with the expanded ids of a search for every base point
    over graph adj.  The first pass searches the k-NN graph; each further
    pass searches the graph of the pass before, so its candidates follow the
    long-range edges that pass created.  Candidates of a point = its k-NN,
    its search pool and the nodes its search expanded (NSG's search-path
    candidates), plus its edges from the pass before, so that edges which
    served navigation compete again instead of being lost.  prune_fn(knn,
    extra, R, ep, alpha) -> (adj, attached) prunes and connects (default:
    the CPU builder, ss_build_graph).  Returns (adj [num][R], ep, info)."""
    import time
    t0 = time.time()
    knn = knn_fn(knn_k)
    t1 = time.time()
    ep = medoid(base)
    info = dict(knn_s=t1 - t0, passes=[])
    if prune_fn is None:
        def prune_fn(knn_, extra_, R_, ep_, alpha_):
            return prune_and_connect(base, metric, knn_, extra_, R_, ep_, alpha_, nthreads)
    num = base.shape[0]
    adj = knn
    path_cap = 2 * nav_L
    for p in range(max(1, passes)):
        ta = time.time()
        nav = None
        if nav_L:
            prev = 0 if p == 0 else adj.shape[1]
            nav = np.empty((num, prev + path_cap + nav_L), dtype=np.uint32)
            if prev:
                nav[:, :prev] = adj
            search_fn(adj, ep, nav_L, path_cap, nav[:, prev + path_cap:], nav[:, prev:prev + path_cap])
        tb = time.time()
        adj, fixed = prune_fn(knn, nav, R, ep, alpha)
        del nav
        tc = time.time()
        deg = float((adj != 0xFFFFFFFF).sum(1).mean())
        info["passes"].append(dict(nav_s=tb - ta, prune_s=tc - tb, attached=fixed, mean_degree=deg))
        if log:
            log(f"graph pass {p}: nav {tb - ta:.1f}s prune {tc - tb:.1f}s attached {fixed} "
                f"mean degree {deg:.1f}")
    if log:
        log(f"graph: knn {info['knn_s']:.1f}s total {time.time() - t0:.1f}s")
    return adj, ep, info


def prune_and_connect_gpu(index, base, knn, extra, R, ep, alpha=1.0):
    """prune_and_connect with the occlusion prunes on the GPU of `index`
    (ss_build_graph_gpu; base is the index's base on the host).  Same graph."""
    base = np.ascontiguousarray(base, dtype=np.float32)
    knn = np.ascontiguousarray(knn, dtype=np.uint32)
    ex = None if extra is None else np.ascontiguousarray(extra, dtype=np.uint32)
    adj = np.empty((base.shape[0], R), dtype=np.uint32)
    fixed = ctypes.c_uint32()
    check(lib.ss_build_graph_gpu(index.handle, fptr(base), uptr(knn), knn.shape[1], uptr(ex),
                                 0 if ex is None else ex.shape[1], R, alpha, ep, uptr(adj), ctypes.byref(fixed)))
    return adj, fixed.value


def gpu_build_graph(index, base, R, knn_k, nav_L, alpha=1.0, passes=3, chunk=131072, nthreads=0,
                    log=None, nn_descent_iters=0, knn_chunk=0, gpu_prune=True):
    """build_graph with the GPU providers of an Index over `base`.  The k-NN
    lists are exact (brute force) unless nn_descent_iters > 0, in which case
    they come from that many NN-descent rounds.  knn_chunk > 0 computes the
    exact lists that many points at a time (each point's k nearest other
    points), with a log line per chunk.  gpu_prune: the occlusion prunes run
    on the GPU (ss_build_graph_gpu, the same graph as the CPU builder)."""

    def knn_fn(k):
        if nn_descent_iters:
            ids, _, upd = index.nn_descent(k, nn_descent_iters)
            if log:
                log(f"nn-descent: insertions per round {upd.tolist()}")
            return ids
        if knn_chunk:
            n = base.shape[0]
            out = np.empty((n, k), dtype=np.uint32)
            for s in range(0, n, knn_chunk):
                e = min(n, s + knn_chunk)
                out[s:e] = index.bruteforce(base[s:e], k, exclude=np.arange(s, e, dtype=np.uint32))[0]
                if log:
                    log(f"knn: {e} of {n} points")
            return out
        return index.knn_graph(k)[0]

    n = base.shape[0]
    log_every = max(chunk, (n // 8 // chunk) * chunk) if n > 4 * chunk else 0

    def search_fn(adj, ep, L, path_cap, out_pool, out_path):
        index.set_graph(adj, ep)
        for s in range(0, n, chunk):
            ids, _, _, pa = index.search_path(base[s:s + chunk], L, L, path_cap)
            out_pool[s:s + chunk] = ids
            out_path[s:s + chunk] = pa
            if log and log_every and (s + chunk) % log_every == 0:
                log(f"navigation search: {min(n, s + chunk)} of {n} points")

    def prune_fn(knn, extra, R_, ep, alpha_):
        if gpu_prune:
            return prune_and_connect_gpu(index, base, knn, extra, R_, ep, alpha_)
        return prune_and_connect(base, index.metric, knn, extra, R_, ep, alpha_, nthreads)

    adj, ep, info = build_graph(base, index.metric, R, knn_k, nav_L, knn_fn, search_fn, alpha,
                                passes, nthreads, log, prune_fn=prune_fn)
    index.set_graph(adj, ep)
    return adj, ep, info
