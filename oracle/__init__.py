"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product never imports it.  What each oracle function restates and how it is
pinned is in oracle/oracle.h.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")
NONE = 0xFFFFFFFF
L2 = 0
IP = 1

_u32 = ctypes.c_uint32
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f32p = ctypes.POINTER(ctypes.c_float)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class Neighbor(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint32), ("distance", ctypes.c_float), ("unexplored", ctypes.c_uint8)]


class Graph(ctypes.Structure):
    _fields_ = [("adj", _u32p), ("num", _u32), ("r_pad", _u32), ("ep", _u32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} not built (make -C oracle)")
        L = ctypes.CDLL(LIB)
        L.ora_dist_l2.restype = ctypes.c_float
        L.ora_dist_l2.argtypes = [_f32p, _f32p, _u32]
        L.ora_dist_ip.restype = ctypes.c_float
        L.ora_dist_ip.argtypes = [_f32p, _f32p, _u32]
        L.ora_insert_into_pool.restype = ctypes.c_int
        L.ora_insert_into_pool.argtypes = [ctypes.POINTER(Neighbor), _u32, Neighbor]
        L.ora_search.restype = ctypes.c_int
        L.ora_search.argtypes = [ctypes.POINTER(Graph), _f32p, _u32, ctypes.c_int, _f32p, _u32, _u32,
                                 _u32p, _f32p, _u32p, _u32p, _u8p, _u32p, _u32]
        L.ora_search_batch.restype = ctypes.c_int
        L.ora_search_batch.argtypes = [ctypes.POINTER(Graph), _f32p, _u32, ctypes.c_int, _f32p, _u32,
                                       _u32p, _u32p, _u32, _u32p, _f32p, _u32p, _u32p, ctypes.c_int,
                                       _f32p, _u32p, _u32]
        L.ora_bruteforce.restype = ctypes.c_int
        L.ora_bruteforce.argtypes = [_f32p, _u32, _f32p, _u32, _u32, ctypes.c_int, _u32, _u32p, _u32p,
                                     _f32p, ctypes.c_int]
        L.ora_scatter_search.restype = ctypes.c_int
        L.ora_scatter_search.argtypes = [ctypes.POINTER(Graph), _f32p, _u32, ctypes.c_int, _f32p, _u32,
                                         _u32, _u32, ctypes.c_int, _u32, _u32p, _f32p, _f32p,
                                         ctypes.POINTER(ctypes.c_uint64)]
        L.ora_scatter_entry.restype = _u32
        L.ora_scatter_entry.argtypes = [_u32, _u32, ctypes.c_int, _u32]
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(_f32p)


def _u(a):
    return a.ctypes.data_as(_u32p) if a is not None else None


def dist(metric, a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    fn = lib().ora_dist_ip if metric == IP else lib().ora_dist_l2
    return fn(_f(a), _f(b), a.shape[0])


def _graph(adj, ep):
    adj = np.ascontiguousarray(adj, dtype=np.uint32)
    g = Graph(_u(adj), adj.shape[0], adj.shape[1], ep)
    return g, adj


def search(base, adj, ep, query, K, L, metric=L2):
    """One query; returns (ids [K], dists [K], ndc, hops)."""
    base = np.ascontiguousarray(base, dtype=np.float32)
    q = np.ascontiguousarray(query, dtype=np.float32)
    g, keep = _graph(adj, ep)
    ids = np.empty(K, dtype=np.uint32)
    ds = np.empty(K, dtype=np.float32)
    ndc, hops = _u32(), _u32()
    vis = np.zeros(base.shape[0], dtype=np.uint8)
    rc = lib().ora_search(ctypes.byref(g), _f(base), base.shape[1], metric, _f(q), K, L, _u(ids), _f(ds),
                          ctypes.byref(ndc), ctypes.byref(hops), vis.ctypes.data_as(_u8p), None, 0)
    if rc:
        raise RuntimeError(f"ora_search failed ({rc})")
    return ids, ds, ndc.value, hops.value


def search_batch(base, adj, ep, queries, K, L, metric=L2, nthreads=0, k_stride=None,
                 latency=False, path_cap=0):
    """Many queries (K, L int or per-query); returns ids, dists, ndc, hops
    (then per-query latency in ms when asked, then the expanded-node lists
    [nq][path_cap] when path_cap > 0)."""
    base = np.ascontiguousarray(base, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = Q.shape[0]
    Ka = np.broadcast_to(np.asarray(K, dtype=np.uint32), (nq,)).copy()
    La = np.broadcast_to(np.asarray(L, dtype=np.uint32), (nq,)).copy()
    ks = int(k_stride or (Ka.max() if nq else 1))
    g, keep = _graph(adj, ep)
    ids = np.empty((nq, ks), dtype=np.uint32)
    ds = np.empty((nq, ks), dtype=np.float32)
    ids.fill(NONE)
    ds.fill(np.inf)
    ndc = np.zeros(nq, dtype=np.uint32)
    hops = np.zeros(nq, dtype=np.uint32)
    lat = np.zeros(nq, dtype=np.float32)
    path = np.empty((nq, path_cap), dtype=np.uint32) if path_cap else None
    rc = lib().ora_search_batch(ctypes.byref(g), _f(base), base.shape[1], metric, _f(Q), nq, _u(Ka),
                                _u(La), ks, _u(ids), _f(ds), _u(ndc), _u(hops), nthreads, _f(lat),
                                _u(path), path_cap)
    if rc:
        raise RuntimeError(f"ora_search_batch failed ({rc})")
    out = (ids, ds, ndc, hops)
    if latency:
        out = out + (lat,)
    if path_cap:
        out = out + (path,)
    return out


def bruteforce(base, queries, k, metric=L2, exclude=None, nthreads=0):
    base = np.ascontiguousarray(base, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = Q.shape[0]
    ids = np.empty((nq, k), dtype=np.uint32)
    ds = np.empty((nq, k), dtype=np.float32)
    ex = None if exclude is None else np.ascontiguousarray(exclude, dtype=np.uint32)
    rc = lib().ora_bruteforce(_f(base), base.shape[0], _f(Q), nq, base.shape[1], metric, k, _u(ex),
                              _u(ids), _f(ds), nthreads)
    if rc:
        raise RuntimeError(f"ora_bruteforce failed ({rc})")
    return ids, ds


def scatter_search(base, adj, ep, queries, K, L, nthreads, metric=L2, seed=1, dists=False):
    """ScatterSearch intra-query search, queries one after another; returns
    (ids [nq][K], latency_ms [nq], total_ndc), and the distances [nq][K]
    last when dists is true."""
    base = np.ascontiguousarray(base, dtype=np.float32)
    Q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = Q.shape[0]
    g, keep = _graph(adj, ep)
    ids = np.empty((nq, K), dtype=np.uint32)
    ds = np.empty((nq, K), dtype=np.float32) if dists else None
    lat = np.zeros(nq, dtype=np.float32)
    ndc = ctypes.c_uint64()
    rc = lib().ora_scatter_search(ctypes.byref(g), _f(base), base.shape[1], metric, _f(Q), nq, K, L,
                                  nthreads, seed, _u(ids), _f(ds) if dists else None, _f(lat),
                                  ctypes.byref(ndc))
    if rc:
        raise RuntimeError(f"ora_scatter_search failed ({rc})")
    return (ids, lat, ndc.value, ds) if dists else (ids, lat, ndc.value)


def scatter_entry(seed, qidx, thread, num):
    """The entry point thread `thread` of scatter_search draws for query qidx."""
    return lib().ora_scatter_entry(seed, qidx, thread, num)


def insert_into_pool(pool, K, nn):
    """pool: ctypes array of Neighbor with K valid entries (room for K+1)."""
    return lib().ora_insert_into_pool(pool, K, nn)


def ref_lib():
    """The reference's own compiled distance routines and InsertIntoPool
    (oracle/_ref/libref.so), or None where it was not built."""
    if not os.path.exists(REF_LIB):
        return None
    R = ctypes.CDLL(REF_LIB)
    for n in ("ref_distance_l2sqr_avx", "ref_distance_ip_avx", "ref_distance_l2sqr", "ref_distance_ip"):
        f = getattr(R, n)
        f.restype = ctypes.c_float
        f.argtypes = [_f32p, _f32p, _u32]
    R.ref_insert_sequence.restype = None
    R.ref_insert_sequence.argtypes = [_u32p, _f32p, _u8p, _u32p, _u32, _u32p, _f32p, _u32,
                                      ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]
    return R


def cpu_graph(base, metric, R, knn_k, nav_L, alpha=1.0, passes=3, nthreads=0, log=None):
    """The graph construction pipeline with the oracle as k-NN and search
    provider (CPU only, for tests)."""
    from scattersearch_amd.graph import build_graph

    def knn_fn(k):
        ex = np.arange(base.shape[0], dtype=np.uint32)
        return bruteforce(base, base, k, metric, exclude=ex, nthreads=nthreads)[0]

    def search_fn(adj, ep, L, path_cap, out_pool, out_path):
        r = search_batch(base, adj, ep, base, L, L, metric, nthreads, path_cap=path_cap)
        out_pool[:] = r[0]
        out_path[:] = r[-1]

    return build_graph(base, metric, R, knn_k, nav_L, knn_fn, search_fn, alpha, passes, nthreads, log)
