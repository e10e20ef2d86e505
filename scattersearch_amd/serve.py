"""Continuous-batching serving (ss_serve_* in the C ABI): submit queries as
they arrive, poll results as they finish; an arriving query is claimed by a
free wavefront of the running serving kernel and searched beside the ones
already on the GPU."""
import ctypes

import numpy as np

from . import _lib
from ._lib import QueryStats, ServeConfig, SSError, check, fptr, lib, uptr
from .index import STATS_DTYPE


def _config(k_max, l_max, ring_slots=0, visited_slots=0, waves_per_block=0, blocks=0, streams=0,
            linger_us=0, admit_us=0, standby_waves=0):
    c = ServeConfig()
    c.k_max, c.l_max, c.ring_slots = k_max, l_max, ring_slots
    c.visited_slots, c.waves_per_block, c.blocks = visited_slots, waves_per_block, blocks
    c.streams, c.linger_us, c.admit_us = streams, linger_us, admit_us
    c.standby_waves = standby_waves
    return c


class Server:
    """A serving loop on the index's GPU.  One thread may submit while
    another polls."""

    def __init__(self, index, k_max, l_max, **kw):
        self.index = index
        self.k_max = int(k_max)
        self.cfg = _config(self.k_max, int(l_max), **kw)
        h = ctypes.c_void_p()
        check(lib.ss_serve_start(index.handle, ctypes.byref(self.cfg), ctypes.byref(h)))
        self._h = h

    def submit(self, query, K, L):
        q = np.ascontiguousarray(query, dtype=np.float32)
        t = ctypes.c_uint64()
        check(lib.ss_submit(self._h, fptr(q), int(K), int(L), ctypes.byref(t)))
        return t.value

    def poll(self, max_results=256):
        """Finished requests: (tickets, ids [n][k_max], dists, stats)."""
        tk = np.empty(max_results, dtype=np.uint64)
        ids = np.empty((max_results, self.k_max), dtype=np.uint32)
        ds = np.empty((max_results, self.k_max), dtype=np.float32)
        st = np.zeros(max_results, dtype=STATS_DTYPE)
        n = ctypes.c_uint32()
        check(lib.ss_poll(self._h, max_results, tk.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), uptr(ids),
                          fptr(ds), st.ctypes.data_as(ctypes.POINTER(QueryStats)), ctypes.byref(n)))
        k = n.value
        _check_status(st[:k])
        return tk[:k], ids[:k], ds[:k], st[:k]

    def stop(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            check(lib.ss_serve_stop(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


# ss_serve_trace: where a replayed request's time went (host clock, us since
# the start) and the server's wavefront counts when its result was seen
TRACE_DTYPE = np.dtype([("submit_us", "<f8"), ("claim_us", "<f8"), ("finish_us", "<f8"), ("seen_us", "<f8"),
                        ("waves_alive", "<u4"), ("waves_standby", "<u4")])


def replay(index, queries, K, L, arrival_us, k_max=None, l_max=None, with_trace=False, **kw):
    """Open-loop load test: query i is submitted arrival_us[i] microseconds
    after the start.  Returns (ids [nq][k_max], dists, latency_us, stats),
    and with_trace=True adds each request's timeline (TRACE_DTYPE: submit,
    claim, finish and seen times in microseconds on the host clock since the
    start, and the server's wavefront counts when it was seen)."""
    q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = q.shape[0]
    Ka = np.ascontiguousarray(np.broadcast_to(np.asarray(K, np.uint32), (nq,)))
    La = np.ascontiguousarray(np.broadcast_to(np.asarray(L, np.uint32), (nq,)))
    arr = np.ascontiguousarray(arrival_us, dtype=np.float64)
    k_max = int(k_max or Ka.max())
    l_max = int(l_max or La.max())
    cfg = _config(k_max, l_max, **kw)
    ids = np.empty((nq, k_max), dtype=np.uint32)
    ds = np.empty((nq, k_max), dtype=np.float32)
    lat = np.empty(nq, dtype=np.float64)
    st = np.zeros(nq, dtype=STATS_DTYPE)
    tr = np.zeros(nq, dtype=TRACE_DTYPE) if with_trace else None
    dp = ctypes.POINTER(ctypes.c_double)
    check(lib.ss_serve_replay(index.handle, ctypes.byref(cfg), fptr(q), nq, uptr(Ka), uptr(La),
                              arr.ctypes.data_as(dp), k_max, uptr(ids), fptr(ds), lat.ctypes.data_as(dp),
                              st.ctypes.data_as(ctypes.POINTER(QueryStats)),
                              tr.ctypes.data_as(ctypes.c_void_p) if with_trace else None))
    _check_status(st)
    return (ids, ds, lat, st, tr) if with_trace else (ids, ds, lat, st)


def _check_status(st):
    bad = np.nonzero(st["status"] == _lib.SS_QUERY_FAULT)[0]
    if bad.size:
        raise SSError(_lib.SS_EINTERNAL, f"serving: internal error, request {int(bad[0])} met a node id "
                                         "outside the index (result not valid)")


def poisson_arrivals(n, rate_qps, seed):
    """Arrival times (microseconds) of a Poisson stream: exponential gaps
    from a seeded generator."""
    gaps = np.random.default_rng(seed).exponential(1e6 / rate_qps, size=n)
    return np.cumsum(gaps) - gaps[0]
