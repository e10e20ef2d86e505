"""Files (fbin, ground truth, NSG), graph construction and metrics, on the
CPU."""
import os
import struct

import numpy as np
import pytest

from scattersearch_amd import SSError, io, metrics
from scattersearch_amd.graph import medoid, prune_and_connect, reachable

NONE = 0xFFFFFFFF


def test_fbin_round_trip_and_layout(tmp_path):
    a = np.random.default_rng(0).standard_normal((37, 13)).astype(np.float32)
    p = tmp_path / "x.fbin"
    io.write_fbin(p, a)
    raw = p.read_bytes()
    assert struct.unpack("<II", raw[:8]) == (37, 13)  # src/util.cpp:48-49
    assert len(raw) == 8 + 37 * 13 * 4
    assert np.array_equal(io.read_fbin(p), a)


def test_fbin_errors(tmp_path):
    with pytest.raises(SSError):
        io.read_fbin(tmp_path / "missing.fbin")
    p = tmp_path / "short.fbin"
    p.write_bytes(struct.pack("<II", 10, 4) + b"\0" * 12)
    with pytest.raises(SSError):
        io.read_fbin(p)


def test_groundtruth_round_trip(tmp_path):
    g = np.random.default_rng(1).integers(0, 1000, (9, 100)).astype(np.uint32)
    p = tmp_path / "gt.bin"
    io.write_groundtruth(p, g)
    raw = p.read_bytes()
    assert struct.unpack("<II", raw[:8]) == (9, 100)  # [nq][k], src/util.cpp:85-86
    assert np.array_equal(io.read_groundtruth(p), g)


def test_nsg_round_trip_and_layout(tmp_path):
    adj = np.full((5, 4), NONE, dtype=np.uint32)
    adj[0, :3] = [1, 2, 3]
    adj[1, :1] = [0]
    adj[2, :4] = [0, 1, 3, 4]
    adj[4, :2] = [2, 3]
    p = tmp_path / "g.nsg"
    io.write_nsg(p, adj, 2)
    raw = np.frombuffer(p.read_bytes(), dtype=np.uint32)
    # [width][ep] then [deg][ids...] per node (src/annsearch.cpp:67-79)
    assert raw[:2].tolist() == [4, 2]
    assert raw[2:].tolist() == [3, 1, 2, 3, 1, 0, 4, 0, 1, 3, 4, 0, 2, 2, 3]
    back, ep = io.read_nsg(p)
    assert ep == 2 and np.array_equal(back, adj)
    wide, _ = io.read_nsg(p, r_pad=8)
    assert np.array_equal(wide[:, :4], adj) and (wide[:, 4:] == NONE).all()


def test_nsg_errors(tmp_path):
    p = tmp_path / "bad.nsg"
    p.write_bytes(np.array([4, 0, 3, 1, 2], dtype=np.uint32).tobytes())  # truncated list
    with pytest.raises(SSError):
        io.read_nsg(p)
    p.write_bytes(np.array([4, 9, 1, 0], dtype=np.uint32).tobytes())  # ep out of range
    with pytest.raises(SSError):
        io.read_nsg(p)
    p.write_bytes(np.array([4, 0, 1, 7], dtype=np.uint32).tobytes())  # id >= num
    with pytest.raises(SSError):
        io.read_nsg(p)
    with pytest.raises(SSError):
        io.read_nsg(tmp_path / "none.nsg")


def test_medoid_is_nearest_to_centroid():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((500, 16)).astype(np.float32)
    c = x.astype(np.float64).mean(0)
    assert medoid(x) == int(np.argmin(((x - c) ** 2).sum(1)))


def test_built_graph_properties(small):
    adj, ep = small["adj"], small["ep"]
    num, R = adj.shape
    assert R == 24
    assert reachable(adj, ep) == num
    for n in range(0, num, 7):
        row = adj[n]
        k = int((row != NONE).sum())
        assert (row[:k] != NONE).all() and (row[k:] == NONE).all()  # prefix-filled
        assert len(set(row[:k].tolist())) == k and n not in row[:k].tolist()


def test_mrng_rule_on_a_line():
    """Points on a line: MRNG keeps only the nearest neighbour on each side."""
    x = np.zeros((7, 8), np.float32)
    x[:, 0] = np.arange(7)
    knn = np.array([[j for j in range(7) if j != i] for i in range(7)], dtype=np.uint32)
    adj, fixed = prune_and_connect(x, 0, knn, None, 6, 3)
    for i in range(7):
        row = sorted(int(v) for v in adj[i] if v != NONE)
        assert set(row) == {j for j in (i - 1, i + 1) if 0 <= j < 7}
    assert fixed == 0


def test_connectivity_pass_attaches_islands():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((40, 8)).astype(np.float32)
    b = rng.standard_normal((40, 8)).astype(np.float32) + 50  # far island
    x = np.concatenate([a, b])
    knn = np.zeros((80, 5), np.uint32)
    for i in range(80):
        d = ((x - x[i]) ** 2).sum(1)
        d[i] = np.inf
        knn[i] = np.argsort(d)[:5]
    adj, fixed = prune_and_connect(x, 0, knn, None, 8, 0)
    assert fixed >= 1
    assert reachable(adj, 0) == 80


def test_recall_and_percentiles():
    res = np.array([[1, 2, 3, 4], [5, 6, 7, 8]])
    gt = np.array([[4, 3, 9, 9], [5, 0, 0, 0]])
    assert metrics.recall_per_query(res, gt, 4).tolist() == [0.5, 0.25]
    assert metrics.recall(res, gt, 4) == pytest.approx(0.375)
    v = np.arange(100)[::-1]
    assert metrics.percentile(v, 0.5) == 50 and metrics.percentile(v, 0.99) == 99
    assert metrics.bytes_per_query(10, 2, 128, 32, 10) == 10 * 512 + 2 * 128 + 512 + 80


def test_search_path_marks_expanded_nodes(small):
    import oracle
    ids, ds, ndc, hops, path = oracle.search_batch(small["base"], small["adj"], small["ep"],
                                                    small["queries"][:10], 10, 32, path_cap=200)
    for r in range(10):
        p = path[r]
        n = int((p != NONE).sum())
        assert n == hops[r] and (p[n:] == NONE).all()
        assert len(set(p[:n].tolist())) == n  # no node is expanded twice
