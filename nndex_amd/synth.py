"""Seeded synthetic workloads for the nndex search path (numpy only).

- ``golden_50k``: 50,000 x 128 iid N(0,1) from ``default_rng(42)``, then one
  query and a batch of four drawn from the same generator, in the order of
  notebooks/python_bindings_demo.ipynb:33-37 (the reference's golden case).
- ``gaussian_mixture``: ``centres`` centres ~ N(0,1); each point is
  centre[u] + spread * N(0,1) with u uniform, then unit-normalised (float32).
  Queries come from the same centres with an independent stream.
- ``anisotropic``: iid N(0,1) scaled per dimension by i^-0.5, unit-normalised.

Generation runs in row blocks to bound peak host memory.
"""

from __future__ import annotations

import numpy as np


def golden_50k():
    rng = np.random.default_rng(42)
    matrix = rng.normal(size=(50_000, 128)).astype(np.float32)
    query = rng.normal(size=(128,)).astype(np.float32)
    queries = rng.normal(size=(4, 128)).astype(np.float32)
    return matrix, query, queries


GOLDEN_50K_INDICES = [43576, 14100, 15993, 35409, 38916]
GOLDEN_50K_SCORES = [0.36015188694000244, 0.35316410660743713, 0.33494800329208374,
                     0.33226925134658813, 0.323153018951416]


def _unit(x: np.ndarray) -> np.ndarray:
    n = np.sqrt(np.einsum("ij,ij->i", x, x, dtype=np.float64)).astype(np.float32)
    x /= n[:, None]
    return x


def mixture_centres(dims: int, centres: int, seed: int = 7) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal((centres, dims), dtype=np.float32)


def gaussian_mixture(rows: int, dims: int, centres: np.ndarray, spread: float, seed: int,
                     block: int = 1 << 18) -> np.ndarray:
    rng = np.random.default_rng(seed)
    out = np.empty((rows, dims), np.float32)
    for r0 in range(0, rows, block):
        n = min(block, rows - r0)
        u = rng.integers(0, centres.shape[0], size=n)
        x = rng.standard_normal((n, dims), dtype=np.float32)
        x *= np.float32(spread)
        x += centres[u]
        out[r0:r0 + n] = _unit(x)
    return out


def mixture_workload(rows: int = 1_000_000, dims: int = 384, n_centres: int = 1000,
                     spread: float = 0.5, n_queries: int = 4096, base_seed: int = 42,
                     query_seed: int = 43):
    """The 1M x 384 exact-search workload: base and queries from one mixture."""
    c = mixture_centres(dims, n_centres)
    base = gaussian_mixture(rows, dims, c, spread, base_seed)
    queries = gaussian_mixture(n_queries, dims, c, spread, query_seed)
    return base, queries


def anisotropic(rows: int, dims: int, seed: int, block: int = 1 << 18) -> np.ndarray:
    rng = np.random.default_rng(seed)
    scale = (np.arange(1, dims + 1, dtype=np.float64) ** -0.5).astype(np.float32)
    out = np.empty((rows, dims), np.float32)
    for r0 in range(0, rows, block):
        n = min(block, rows - r0)
        x = rng.standard_normal((n, dims), dtype=np.float32) * scale
        out[r0:r0 + n] = _unit(x)
    return out
