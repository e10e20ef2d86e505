"""fbin / ground-truth / NSG files through the C ABI host functions.

Formats: fbin ``[u32 num][u32 dim][num*dim f32]`` (load_fbin,
src/util.cpp:40-70); ground truth ``[u32 nq][u32 k][nq*k u32]``
(load_groundtruth, src/util.cpp:72-93); NSG ``[u32 width][u32 ep]`` then per
node ``[u32 deg][deg u32]`` (LoadGraph, src/annsearch.cpp:64-80).
"""
import ctypes

import numpy as np

from ._lib import check, fptr, lib, uptr


def _u32():
    return ctypes.c_uint32()


def read_fbin(path):
    n, d = _u32(), _u32()
    check(lib.ss_fbin_header(str(path).encode(), ctypes.byref(n), ctypes.byref(d)))
    a = np.empty((n.value, d.value), dtype=np.float32)
    check(lib.ss_fbin_read(str(path).encode(), fptr(a), n.value, d.value))
    return a


def write_fbin(path, data):
    a = np.ascontiguousarray(data, dtype=np.float32)
    check(lib.ss_fbin_write(str(path).encode(), fptr(a), a.shape[0], a.shape[1]))


def read_groundtruth(path):
    n, k = _u32(), _u32()
    check(lib.ss_gt_header(str(path).encode(), ctypes.byref(n), ctypes.byref(k)))
    a = np.empty((n.value, k.value), dtype=np.uint32)
    check(lib.ss_gt_read(str(path).encode(), uptr(a), n.value, k.value))
    return a


def write_groundtruth(path, ids):
    a = np.ascontiguousarray(ids, dtype=np.uint32)
    check(lib.ss_gt_write(str(path).encode(), uptr(a), a.shape[0], a.shape[1]))


def read_nsg(path, r_pad=None):
    """Returns (adj [num][r_pad] padded with SS_NONE, ep)."""
    n, md, ep = _u32(), _u32(), _u32()
    check(lib.ss_nsg_header(str(path).encode(), ctypes.byref(n), ctypes.byref(md), ctypes.byref(ep)))
    width = int(r_pad) if r_pad else max(md.value, 1)
    adj = np.empty((n.value, width), dtype=np.uint32)
    check(lib.ss_nsg_read(str(path).encode(), uptr(adj), n.value, width))
    return adj, ep.value


def write_nsg(path, adj, ep):
    a = np.ascontiguousarray(adj, dtype=np.uint32)
    check(lib.ss_nsg_write(str(path).encode(), uptr(a), a.shape[0], a.shape[1], int(ep)))
