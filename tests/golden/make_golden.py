#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/.

ref_vectors.npz   inputs and outputs of the reference's own compiled
                  distance_l2sqr_avx / distance_ip_avx (src/distance.cpp)
                  and of its InsertIntoPool (include/util.h:41-78) driven
                  the way the search loop drives it; produced through
                  oracle/_ref/libref.so, which `make -C oracle ref` builds
                  from the reference sources.  Needs that library.
tiny_search.npz   a 2,000 x 32 mixture, its R=16 graph and 50 queries with
                  the oracle's ids, distances, distance counts and hops at
                  L=32, K=10: a fixed instance that pins the oracle (and,
                  in the GPU tests, the kernel) against drift of the data
                  generator or graph builder.

Run from the repository root: python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def ref_vectors():
    R = oracle.ref_lib()
    if R is None:
        print("oracle/_ref/libref.so missing: run make -C oracle ref", file=sys.stderr)
        return
    rng = np.random.default_rng(20261203)
    fp = ctypes.POINTER(ctypes.c_float)
    out = {}
    for dim in (96, 128, 200):
        a = (rng.standard_normal((32, dim)) * rng.uniform(0.1, 10, (32, 1))).astype(np.float32)
        b = (rng.standard_normal((32, dim)) * rng.uniform(0.1, 10, (32, 1))).astype(np.float32)
        l2 = np.empty(32, np.float32)
        ip = np.empty(32, np.float32)
        for i in range(32):
            pa, pb = a[i].ctypes.data_as(fp), b[i].ctypes.data_as(fp)
            l2[i] = R.ref_distance_l2sqr_avx(pa, pb, dim)
            ip[i] = R.ref_distance_ip_avx(pa, pb, dim)
        out[f"a{dim}"], out[f"b{dim}"], out[f"l2_{dim}"], out[f"ip_{dim}"] = a, b, l2, ip
    # InsertIntoPool sequences: pool of L+1, t valid entries, then n inserts
    # with the search loop's tail filter.
    u32p = ctypes.POINTER(ctypes.c_uint32)
    seqs = []
    for s in range(24):
        L = int(rng.integers(2, 33))
        t = int(rng.integers(1, L + 1))
        pd = np.sort(np.round(rng.uniform(0, 10, L + 1) * 2) / 2).astype(np.float32)
        pi = rng.permutation(500)[:L + 1].astype(np.uint32)
        n = 40
        ds = (np.round(rng.uniform(-1, 11, n) * 2) / 2).astype(np.float32)
        ids = rng.integers(0, 500, n).astype(np.uint32)
        r_ids, r_d, r_u = pi.copy(), pd.copy(), np.ones(L + 1, np.uint8)
        rt = ctypes.c_uint32(t)
        pos = np.empty(n, np.int32)
        R.ref_insert_sequence(r_ids.ctypes.data_as(u32p), r_d.ctypes.data_as(fp),
                              r_u.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(rt), L,
                              ids.ctypes.data_as(u32p), ds.ctypes.data_as(fp), n,
                              pos.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 1)
        seqs.append((L, t, pi, pd, ids, ds, pos, r_ids, r_d, rt.value))
    for i, (L, t, pi, pd, ids, ds, pos, r_ids, r_d, t1) in enumerate(seqs):
        out[f"ins{i}_meta"] = np.array([L, t, t1], np.int64)
        out[f"ins{i}_pool_ids"], out[f"ins{i}_pool_d"] = pi, pd
        out[f"ins{i}_ids"], out[f"ins{i}_d"] = ids, ds
        out[f"ins{i}_pos"], out[f"ins{i}_out_ids"], out[f"ins{i}_out_d"] = pos, r_ids, r_d
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote ref_vectors.npz")


def tiny_search():
    from scattersearch_amd import data
    base, q = data.mixture(2000, 50, 32, 6, 0.3, seed=424242)
    adj, ep, _ = oracle.cpu_graph(base, 0, 16, 16, 24, passes=3)
    ids, ds, ndc, hops = oracle.search_batch(base, adj, ep, q, 10, 32)
    gt, gd = oracle.bruteforce(base, q, 10)
    np.savez_compressed(os.path.join(HERE, "tiny_search.npz"), base=base, queries=q, adj=adj,
                        ep=np.array(ep), K=np.array(10), L=np.array(32), ids=ids, dists=ds, ndc=ndc,
                        hops=hops, gt=gt, gt_dists=gd)
    print("wrote tiny_search.npz")


if __name__ == "__main__":
    ref_vectors()
    tiny_search()
