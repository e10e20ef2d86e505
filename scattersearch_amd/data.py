"""Seeded synthetic workloads (Gaussian mixtures) for the query path.

Points are drawn as centre[c] + sigma * N(0, I) with the component c uniform
over the centres, centres ~ N(0, I).  Each workload has one seed; centres,
base points and queries come from three independent child streams of it
(numpy SeedSequence.spawn), so a workload is reproducible on any machine with
the same numpy.
"""
from dataclasses import dataclass, field

import numpy as np

from .index import METRIC_INNER_PRODUCT, METRIC_L2

SEED_BASE = 20261203


@dataclass(frozen=True)
class Workload:
    name: str
    num: int
    dim: int
    nq: int
    centres: int
    sigma: float
    metric: int
    R: int
    knn_k: int
    K: int
    L_list: tuple
    seed: int
    lognormal_sigma: float = 0.0
    unit_norm: bool = False
    nav_L: int = 64
    extra: dict = field(default_factory=dict)


WORKLOADS = {
    # 20k x 128, 50 components, sigma 0.3, R=32, L=64, K=10 (CPU-runnable case)
    "c0": Workload("mixture-20k-d128-R32", 20_000, 128, 200, 50, 0.3, METRIC_L2, 32, 32, 10,
                   (64,), SEED_BASE + 0),
    # 1M x 128, 1000 components, sigma 0.25, 10k queries, L sweep, K=10
    "c1": Workload("mixture-1M-d128-R32", 1_000_000, 128, 10_000, 1000, 0.25, METRIC_L2, 32, 32,
                   10, (32, 64, 128, 256), SEED_BASE + 1),
    # 1.2M x 200 inner product, log-normal norms then unit-normalised, R=64
    "c2": Workload("mixture-1.2M-d200-ip-R64", 1_200_000, 200, 10_000, 1000, 0.25,
                   METRIC_INNER_PRODUCT, 64, 64, 10, (), SEED_BASE + 2, lognormal_sigma=0.5,
                   unit_norm=True),
    # 10M x 96, 4096 components, sigma 0.2, 100k queries, k-NN by 10 NN-descent rounds
    "c3": Workload("mixture-10M-d96-R32", 10_000_000, 96, 100_000, 4096, 0.2, METRIC_L2, 32, 32, 10,
                   (64, 128), SEED_BASE + 3, extra={"nn_descent_iters": 10}),
}


def mixture(num, nq, dim, centres, sigma, seed, lognormal_sigma=0.0, unit_norm=False,
            chunk=1 << 18):
    """Returns (base [num][dim] f32, queries [nq][dim] f32)."""
    ss = np.random.SeedSequence(seed)
    c_seed, b_seed, q_seed = ss.spawn(3)
    C = np.random.default_rng(c_seed).standard_normal((centres, dim), dtype=np.float32)

    def draw(seed_seq, n):
        rng = np.random.default_rng(seed_seq)
        out = np.empty((n, dim), dtype=np.float32)
        for s in range(0, n, chunk):
            e = min(n, s + chunk)
            assign = rng.integers(0, centres, size=e - s)
            x = rng.standard_normal((e - s, dim), dtype=np.float32)
            x *= np.float32(sigma)
            x += C[assign]
            if lognormal_sigma:
                x *= rng.lognormal(0.0, lognormal_sigma, size=e - s).astype(np.float32)[:, None]
            if unit_norm:
                x /= np.linalg.norm(x, axis=1, keepdims=True).astype(np.float32)
            out[s:e] = x
        return out

    return draw(b_seed, num), draw(q_seed, nq)


def generate(w: Workload, num=None, nq=None):
    """Base and queries of a workload (optionally a prefix-sized variant)."""
    return mixture(num or w.num, nq or w.nq, w.dim, w.centres, w.sigma, w.seed,
                   w.lognormal_sigma, w.unit_norm)
