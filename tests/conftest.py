import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def c0():
    """The 20k x 128 mixture workload with its graph built on the CPU."""
    import oracle
    from scattersearch_amd import data
    w = data.WORKLOADS["c0"]
    base, queries = data.generate(w)
    adj, ep, _ = oracle.cpu_graph(base, w.metric, w.R, w.knn_k, w.nav_L, passes=1)
    return dict(w=w, base=base, queries=queries, adj=adj, ep=ep)


@pytest.fixture(scope="session")
def small():
    """3000 x 128 mixture with a CPU-built graph (R=24)."""
    import oracle
    from scattersearch_amd import data
    base, queries = data.mixture(3000, 64, 128, 8, 0.3, seed=7)
    adj, ep, _ = oracle.cpu_graph(base, 0, 24, 24, 32, passes=1)
    return dict(base=base, queries=queries, adj=adj, ep=ep)


def random_graph(rng, num, r_pad, min_deg=1):
    """Random prefix-filled adjacency without repeats inside a row."""
    adj = np.full((num, r_pad), 0xFFFFFFFF, dtype=np.uint32)
    for n in range(num):
        d = int(rng.integers(min_deg, r_pad + 1))
        adj[n, :d] = rng.choice(num, size=min(d, num), replace=False)
    return adj
