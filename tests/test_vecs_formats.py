"""SIFT-style vecs files (load_fvecs, src/util.cpp:6-38) and the dataset
conversions of scripts/splitfbin.py and scripts/hdf5tofbin.py, on the CPU.
Layouts are checked against bytes written with struct."""
import struct

import numpy as np
import pytest

from scattersearch_amd import SSError, io


def _fvecs_bytes(a):
    return b"".join(struct.pack("<i", a.shape[1]) + r.astype("<f4").tobytes() for r in a)


def test_fvecs_read_matches_struct_layout(tmp_path):
    a = np.random.default_rng(3).standard_normal((23, 17)).astype(np.float32)
    p = tmp_path / "x.fvecs"
    p.write_bytes(_fvecs_bytes(a))
    assert np.array_equal(io.read_fvecs(p), a)


def test_fvecs_write_layout_and_round_trip(tmp_path):
    a = np.random.default_rng(4).standard_normal((9, 128)).astype(np.float32)
    p = tmp_path / "y.fvecs"
    io.write_fvecs(p, a)
    assert p.read_bytes() == _fvecs_bytes(a)
    assert np.array_equal(io.read_fvecs(p), a)


def test_ivecs_and_bvecs(tmp_path):
    g = np.random.default_rng(5).integers(0, 1 << 20, (11, 100)).astype(np.uint32)
    p = tmp_path / "gt.ivecs"
    io.write_ivecs(p, g)
    raw = p.read_bytes()
    assert struct.unpack("<i", raw[:4])[0] == 100 and len(raw) == 11 * (4 + 400)
    assert np.array_equal(io.read_ivecs(p), g)
    b = np.random.default_rng(6).integers(0, 256, (7, 96)).astype(np.uint8)
    pb = tmp_path / "x.bvecs"
    pb.write_bytes(b"".join(struct.pack("<i", 96) + r.tobytes() for r in b))
    assert np.array_equal(io.read_bvecs(pb), b.astype(np.float32))


def test_vecs_errors(tmp_path):
    with pytest.raises(SSError):
        io.read_fvecs(tmp_path / "missing.fvecs")
    a = np.ones((3, 8), dtype=np.float32)
    p = tmp_path / "trunc.fvecs"
    p.write_bytes(_fvecs_bytes(a)[:-4])  # not a whole number of records
    with pytest.raises(SSError):
        io.read_fvecs(p)
    q = tmp_path / "mixed.fvecs"  # same size per record as dim 8, but one record claims dim 7
    raw = bytearray(_fvecs_bytes(a))
    raw[36:40] = struct.pack("<i", 7)
    q.write_bytes(bytes(raw))
    with pytest.raises(SSError):
        io.read_fvecs(q)
    e = tmp_path / "empty.fvecs"
    e.write_bytes(b"")
    assert io.read_fvecs(e).shape == (0, 0)


def test_conversions_to_fbin_and_groundtruth(tmp_path):
    a = np.random.default_rng(7).standard_normal((50, 16)).astype(np.float32)
    io.write_fvecs(tmp_path / "base.fvecs", a)
    io.fvecs_to_fbin(tmp_path / "base.fvecs", tmp_path / "base.fbin")
    assert np.array_equal(io.read_fbin(tmp_path / "base.fbin"), a)
    g = np.arange(50 * 10, dtype=np.uint32).reshape(50, 10)
    io.write_ivecs(tmp_path / "gt.ivecs", g)
    io.ivecs_to_groundtruth(tmp_path / "gt.ivecs", tmp_path / "gt.bin")
    assert np.array_equal(io.read_groundtruth(tmp_path / "gt.bin"), g)


def test_split_fbin_takes_queries_from_the_end(tmp_path):
    a = np.random.default_rng(8).standard_normal((40, 8)).astype(np.float32)
    io.write_fbin(tmp_path / "all.fbin", a)
    io.split_fbin(tmp_path / "all.fbin", tmp_path / "b.fbin", tmp_path / "q.fbin", num_queries=12)
    assert np.array_equal(io.read_fbin(tmp_path / "b.fbin"), a[:28])
    assert np.array_equal(io.read_fbin(tmp_path / "q.fbin"), a[28:])
    with pytest.raises(ValueError):
        io.split_fbin(tmp_path / "all.fbin", tmp_path / "b2.fbin", tmp_path / "q2.fbin", num_queries=41)


def test_hdf5_conversion_needs_h5py(tmp_path):
    try:
        import h5py  # noqa: F401
    except ImportError:
        with pytest.raises(RuntimeError):
            io.hdf5_to_fbin(tmp_path / "x.hdf5", base_dst=tmp_path / "b.fbin")
        pytest.skip("h5py is not installed: the HDF5 path is only checked to fail clearly")
    import h5py
    tr = np.random.default_rng(9).standard_normal((30, 8))
    te = np.random.default_rng(10).standard_normal((5, 8))
    nb = np.random.default_rng(11).integers(0, 30, (5, 100)).astype(np.int64)
    with h5py.File(tmp_path / "d.hdf5", "w") as f:
        f["train"], f["test"], f["neighbors"] = tr, te, nb
    io.hdf5_to_fbin(tmp_path / "d.hdf5", tmp_path / "b.fbin", tmp_path / "q.fbin", tmp_path / "g.bin")
    assert np.array_equal(io.read_fbin(tmp_path / "b.fbin"), tr.astype(np.float32))
    assert np.array_equal(io.read_fbin(tmp_path / "q.fbin"), te.astype(np.float32))
    assert np.array_equal(io.read_groundtruth(tmp_path / "g.bin"), nb.astype(np.uint32))
