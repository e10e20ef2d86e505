"""fbin / ground-truth / NSG / vecs files through the C ABI host functions,
and the dataset conversions of the reference's scripts.

Formats: fbin ``[u32 num][u32 dim][num*dim f32]`` (load_fbin,
src/util.cpp:40-70); ground truth ``[u32 nq][u32 k][nq*k u32]``
(load_groundtruth, src/util.cpp:72-93); NSG ``[u32 width][u32 ep]`` then per
node ``[u32 deg][deg u32]`` (LoadGraph, src/annsearch.cpp:64-80); fvecs /
ivecs / bvecs records ``[i32 dim][dim f32 | i32 | u8]`` (load_fvecs,
src/util.cpp:6-38).
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


def _vecs_shape(path, elem):
    n, d = _u32(), _u32()
    check(lib.ss_vecs_header(str(path).encode(), elem, ctypes.byref(n), ctypes.byref(d)))
    return n.value, d.value


def read_fvecs(path):
    n, d = _vecs_shape(path, 4)
    a = np.empty((n, d), dtype=np.float32)
    check(lib.ss_fvecs_read(str(path).encode(), fptr(a), n, d))
    return a


def read_ivecs(path):
    n, d = _vecs_shape(path, 4)
    a = np.empty((n, d), dtype=np.uint32)
    check(lib.ss_ivecs_read(str(path).encode(), uptr(a), n, d))
    return a


def read_bvecs(path):
    """Byte vectors widened to float32."""
    n, d = _vecs_shape(path, 1)
    a = np.empty((n, d), dtype=np.float32)
    check(lib.ss_bvecs_read(str(path).encode(), fptr(a), n, d))
    return a


def write_fvecs(path, data):
    a = np.ascontiguousarray(data, dtype=np.float32)
    check(lib.ss_fvecs_write(str(path).encode(), fptr(a), a.shape[0], a.shape[1]))


def write_ivecs(path, data):
    a = np.ascontiguousarray(data, dtype=np.uint32)
    check(lib.ss_ivecs_write(str(path).encode(), uptr(a), a.shape[0], a.shape[1]))


def fvecs_to_fbin(src, dst):
    """SIFT-style base or query file to the reference's fbin input."""
    write_fbin(dst, read_fvecs(src))


def ivecs_to_groundtruth(src, dst):
    """SIFT-style ground truth (ivecs) to the reference's ground-truth file."""
    write_groundtruth(dst, read_ivecs(src))


def split_fbin(src, base_dst, query_dst, num_queries=1000):
    """The last num_queries vectors become the query set, the rest the base
    (scripts/splitfbin.py)."""
    a = read_fbin(src)
    if a.shape[0] < num_queries:
        raise ValueError(f"{src} holds {a.shape[0]} vectors, fewer than {num_queries}")
    write_fbin(base_dst, a[:a.shape[0] - num_queries])
    write_fbin(query_dst, a[a.shape[0] - num_queries:])


def hdf5_to_fbin(src, base_dst=None, query_dst=None, gt_dst=None):
    """ann-benchmarks HDF5 (datasets train / test / neighbors) to fbin base,
    fbin queries and ground truth (scripts/hdf5tofbin.py).  float64 data is
    narrowed to float32, int64 neighbour ids to uint32.  Needs h5py."""
    try:
        import h5py
    except ImportError as exc:
        raise RuntimeError("hdf5_to_fbin needs the h5py package, which is not installed") from exc
    with h5py.File(src, "r") as f:
        if base_dst:
            write_fbin(base_dst, np.asarray(f["train"][:], dtype=np.float32))
        if query_dst:
            write_fbin(query_dst, np.asarray(f["test"][:], dtype=np.float32))
        if gt_dst:
            nb = np.asarray(f["neighbors"][:])
            if nb.dtype not in (np.int32, np.uint32, np.int64):
                raise ValueError(f"unsupported neighbour id type {nb.dtype}")
            write_groundtruth(gt_dst, nb.astype(np.uint32))
