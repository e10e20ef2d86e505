"""Python view of the C ABI: an Index owns one ss_index (base vectors and
graph resident in HBM) and runs searches and brute force on the GPU.

``ANNSearch`` below keeps the reference's method names and argument meaning
(include/annsearch.h:37-84) on top of it.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import SS_NONE, SearchParams, QueryStats, check, fptr, uptr, lib

METRIC_L2 = 0
METRIC_INNER_PRODUCT = 1

STATS_DTYPE = np.dtype([("ndc", np.uint32), ("hops", np.uint32),
                        ("lossy", np.uint32), ("status", np.uint32),
                        ("t_start", np.uint64), ("t_end", np.uint64)])


def dim_padded(dim):
    return int(lib.ss_dim_padded(dim))


def _params(mode=_lib.SS_MODE_BATCHED, waves_per_block=0, visited_slots=0,
            blocks=0, use_graph=False, expand_width=0, waves_per_query=0, query_order=0,
            team_size=0, pool_layout=0, pool_classes=0, visited_bucket=0):
    p = SearchParams()
    p.visited_bucket = visited_bucket
    p.query_order = query_order
    p.pool_classes = pool_classes
    p.team_size = team_size
    p.pool_layout = pool_layout
    p.expand_width = expand_width
    p.waves_per_query = waves_per_query
    p.mode = mode
    p.waves_per_block = waves_per_block
    p.visited_slots = visited_slots
    p.blocks = blocks
    p.use_graph = 1 if use_graph else 0
    return p


class Index:
    """Base vectors (and, once loaded, a graph) resident on the current GPU.

    ANNSearch::ANNSearch (src/annsearch.cpp:11-58) borrows the caller's
    array; this copies it to HBM.
    """

    def __init__(self, base, metric=METRIC_L2):
        base = np.ascontiguousarray(base, dtype=np.float32)
        if base.ndim != 2:
            raise ValueError("base must be [num][dim]")
        self.num, self.dim = base.shape
        self.metric = int(metric)
        self.dim_pad = dim_padded(self.dim)
        h = ctypes.c_void_p()
        check(lib.ss_create(self.dim, self.num, fptr(base), self.metric, ctypes.byref(h)))
        self._h = h

    @classmethod
    def from_device(cls, d_base_ptr, num, dim, metric=METRIC_L2):
        """Index over an existing device array [num][dim_pad] (copied)."""
        self = cls.__new__(cls)
        self.num, self.dim, self.metric = int(num), int(dim), int(metric)
        self.dim_pad = dim_padded(dim)
        h = ctypes.c_void_p()
        check(lib.ss_create_from_device(self.dim, self.num, ctypes.c_void_p(d_base_ptr),
                                        self.metric, ctypes.byref(h)))
        self._h = h
        return self

    def close(self):
        if getattr(self, "_h", None):
            lib.ss_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self):
        d, n, r, e = (ctypes.c_uint32() for _ in range(4))
        m = ctypes.c_int()
        check(lib.ss_index_info(self._h, ctypes.byref(d), ctypes.byref(n), ctypes.byref(r),
                                ctypes.byref(e), ctypes.byref(m)))
        return dict(dim=d.value, num=n.value, r_pad=r.value, ep=e.value, metric=m.value)

    def wall_clock_khz(self):
        k = ctypes.c_uint32()
        check(lib.ss_wall_clock_khz(self._h, ctypes.byref(k)))
        return k.value

    # -- graph ------------------------------------------------------------
    def load_graph(self, path):
        check(lib.ss_load_graph(self._h, str(path).encode()))

    def set_graph(self, adj, ep):
        adj = np.ascontiguousarray(adj, dtype=np.uint32)
        if adj.ndim != 2 or adj.shape[0] != self.num:
            raise ValueError("adjacency must be [num][r_pad]")
        check(lib.ss_set_graph(self._h, uptr(adj), adj.shape[1], int(ep)))

    def get_graph(self):
        info = self.info()
        adj = np.empty((self.num, info["r_pad"]), dtype=np.uint32)
        r, e = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib.ss_get_graph(self._h, uptr(adj), ctypes.byref(r), ctypes.byref(e)))
        return adj, e.value

    # -- search -----------------------------------------------------------
    def search(self, queries, K, L, k_stride=None, mode=_lib.SS_MODE_BATCHED,
               waves_per_block=0, visited_slots=0, blocks=0, use_graph=False, expand_width=0,
               waves_per_query=0, query_order=0, team_size=0, pool_layout=0, pool_classes=0,
               visited_bucket=0):
        """Search a batch of host queries.  K and L are ints or per-query
        arrays.  query_order is SS_ORDER_* (scheduling only; results do not
        depend on it); team_size (4, 8), pool_layout (SS_POOL_*),
        pool_classes (SS_CLASSES_*) and visited_bucket (8 or 16, optionally
        | SS_VISITED_TWO_CHOICE) do not change results either.  Returns (ids [nq][k_stride], dists, stats)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 1:
            q = q[None, :]
        nq = q.shape[0]
        if q.shape[1] != self.dim:
            raise ValueError("query dimension differs from the base")
        Ka = La = None
        k_def = l_def = 0
        if np.ndim(K) == 0:
            k_def = int(K)
        else:
            Ka = np.ascontiguousarray(K, dtype=np.uint32)
        if np.ndim(L) == 0:
            l_def = int(L)
        else:
            La = np.ascontiguousarray(L, dtype=np.uint32)
        kmax = int(Ka.max()) if Ka is not None and nq else k_def
        ks = int(k_stride) if k_stride else max(kmax, 1)
        ids = np.empty((nq, ks), dtype=np.uint32)
        dists = np.empty((nq, ks), dtype=np.float32)
        stats = np.zeros(nq, dtype=STATS_DTYPE)
        p = _params(mode, waves_per_block, visited_slots, blocks, use_graph, expand_width, waves_per_query,
                    query_order, team_size, pool_layout, pool_classes, visited_bucket)
        check(lib.ss_search(self._h, fptr(q), nq, uptr(Ka), uptr(La), k_def, l_def, ks,
                            ctypes.byref(p), uptr(ids), fptr(dists),
                            stats.ctypes.data_as(ctypes.POINTER(QueryStats))))
        return ids, dists, stats

    def search_path(self, queries, K, L, path_cap, visited_slots=0):
        """Search with one K and L, also returning the expanded node ids of
        every query in order ([nq][path_cap], SS_NONE-padded)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, K), dtype=np.uint32)
        dists = np.empty((nq, K), dtype=np.float32)
        stats = np.zeros(nq, dtype=STATS_DTYPE)
        path = np.empty((nq, path_cap), dtype=np.uint32)
        p = _params(visited_slots=visited_slots)
        check(lib.ss_search_path(self._h, fptr(q), nq, int(K), int(L), int(K), ctypes.byref(p),
                                 int(path_cap), uptr(ids), fptr(dists),
                                 stats.ctypes.data_as(ctypes.POINTER(QueryStats)), uptr(path)))
        return ids, dists, stats, path

    def search_device(self, d_queries, nq, k_max, l_max, k_stride, d_out_ids, d_out_dists,
                      d_stats=0, d_K=0, d_L=0, stream=0, mode=_lib.SS_MODE_BATCHED,
                      waves_per_block=0, visited_slots=0, blocks=0, use_graph=False, expand_width=0,
                      waves_per_query=0, query_order=0, team_size=0, pool_layout=0, pool_classes=0,
                      visited_bucket=0):
        """Enqueue a search on device buffers (raw pointers, e.g. from torch
        tensors' data_ptr()); queries are [nq][dim_pad]."""
        p = _params(mode, waves_per_block, visited_slots, blocks, use_graph, expand_width, waves_per_query,
                    query_order, team_size, pool_layout, pool_classes, visited_bucket)
        vp = ctypes.c_void_p
        check(lib.ss_search_device(self._h, vp(d_queries), nq, vp(d_K or None), vp(d_L or None),
                                   k_max, l_max, k_stride, ctypes.byref(p), vp(d_out_ids),
                                   vp(d_out_dists or None), vp(d_stats or None),
                                   vp(stream or None)))

    def autotune(self, queries, K, L, max_recall_loss=0.0):
        """Pick and store launch parameters for pool size L (see ss_autotune).
        Returns dict(expand_width, waves_per_block, visited_slots, ms, recall)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        p = SearchParams()
        ms, rec = ctypes.c_float(), ctypes.c_float()
        check(lib.ss_autotune(self._h, fptr(q), q.shape[0], int(K), int(L), float(max_recall_loss),
                              ctypes.byref(p), ctypes.byref(ms), ctypes.byref(rec)))
        return dict(expand_width=p.expand_width, waves_per_block=p.waves_per_block,
                    visited_slots=p.visited_slots, waves_per_query=p.waves_per_query,
                    team_size=p.team_size, pool_layout=p.pool_layout, visited_bucket=p.visited_bucket,
                    ms=ms.value, recall=rec.value)

    def bruteforce(self, queries, k, exclude=None):
        """Exact top-k by (distance, id) on the GPU."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), dtype=np.uint32)
        dists = np.empty((nq, k), dtype=np.float32)
        ex = None if exclude is None else np.ascontiguousarray(exclude, dtype=np.uint32)
        check(lib.ss_bruteforce(self._h, fptr(q), nq, k, uptr(ex), uptr(ids), fptr(dists)))
        return ids, dists

    def bruteforce_device(self, d_queries, nq, k, d_out_ids, d_out_dists, exclude_self=False,
                          query_base_offset=0, stream=0):
        vp = ctypes.c_void_p
        check(lib.ss_bruteforce_device(self._h, vp(d_queries), nq, k, 1 if exclude_self else 0,
                                       query_base_offset, vp(d_out_ids), vp(d_out_dists),
                                       vp(stream or None)))

    def knn_graph(self, k):
        """Each base point's k nearest other base points (GPU)."""
        ids = np.empty((self.num, k), dtype=np.uint32)
        dists = np.empty((self.num, k), dtype=np.float32)
        check(lib.ss_knn_graph(self._h, k, uptr(ids), fptr(dists)))
        return ids, dists

    def nn_descent(self, k, iters=10, seed=20261203):
        """Approximate k-NN lists by NN-descent on the GPU; returns (ids,
        dists, insertions per iteration)."""
        ids = np.empty((self.num, k), dtype=np.uint32)
        dists = np.empty((self.num, k), dtype=np.float32)
        upd = np.zeros(max(iters, 1), dtype=np.uint32)
        check(lib.ss_nn_descent(self._h, k, iters, seed, uptr(ids), fptr(dists), uptr(upd)))
        return ids, dists, upd[:iters]


class Metric:
    """enum Metric (include/annsearch.h:22-26)."""
    L2 = METRIC_L2
    INNER_PRODUCT = METRIC_INNER_PRODUCT


class Neighbor:
    """Neighbor (include/util.h:23-34): id and distance of a result."""
    __slots__ = ("id", "distance", "unexplored")

    def __init__(self, id, distance, unexplored=False):
        self.id = int(id)
        self.distance = float(distance)
        self.unexplored = bool(unexplored)

    def __lt__(self, other):
        return self.distance < other.distance

    def __repr__(self):
        return f"Neighbor({self.id}, {self.distance})"


class ANNSearch:
    """The reference engine's interface (include/annsearch.h:37-84) over the
    GPU index.  Method names and argument meaning follow the reference; the
    visited bitset argument is accepted and ignored (the visited state lives
    on the device and is private to each query)."""

    def __init__(self, dim, num, base, m=Metric.L2):
        base = np.asarray(base, dtype=np.float32).reshape(num, dim)
        self.dimension = dim
        self.base_num = num
        self.index = Index(base, m)
        self.groundtruth = []
        self.default_ep = 0
        self.dist_comps = 0
        self.hop_count = 0

    def LoadGraph(self, filename):
        """src/annsearch.cpp:64-80"""
        self.index.load_graph(filename)
        self.default_ep = self.index.info()["ep"]

    def LoadGroundtruth(self, filename):
        """src/annsearch.cpp:82-101"""
        from .io import read_groundtruth
        gt = read_groundtruth(filename)
        self.groundtruth = [list(r) for r in gt]

    def SearchArraySimulation(self, query, query_id, K, L, flags=None, indices=None):
        """src/annsearch.cpp:182-259: returns the K ids (also written into
        `indices` when a list is given, like the reference's out-parameter)."""
        ids, _, st = self.index.search(np.asarray(query, dtype=np.float32)[None, :], int(K), int(L))
        self.dist_comps += int(st["ndc"][0])
        self.hop_count += int(st["hops"][0])
        out = [int(x) for x in ids[0, :K]]
        if indices is not None:
            indices[:] = out
        return out

    def SearchArraySimulationForPipeline(self, query, query_id, thread_id, num_threads, K, L,
                                         flags=None, indices=None):
        """src/annsearch.cpp:261-324 with thread_id=0, num_threads=1: the
        K nearest as Neighbor(id, distance)."""
        if num_threads != 1 or thread_id != 0:
            raise ValueError("only the single-slice form (thread 0 of 1) runs on the GPU")
        ids, d, st = self.index.search(np.asarray(query, dtype=np.float32)[None, :], int(K), int(L))
        self.dist_comps += int(st["ndc"][0])
        out = [Neighbor(i, x) for i, x in zip(ids[0, :K], d[0, :K])]
        if indices is not None:
            indices[:] = out
        return out

    def SearchBatch(self, queries, K, L, **kw):
        """All queries in one persistent launch: (ids, dists, stats)."""
        return self.index.search(queries, K, L, **kw)
