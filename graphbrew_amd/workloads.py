"""Synthetic workloads shared by bench.py and the GPU tests.

many_small_graphs: the batch of small seeded R-MAT graphs of uneven size that
the many-graphs entry points (gbk_cc_many, gbk_pr_many) are measured on —
scales 10-16, edge factors 8, 16 or 32, quadrant probabilities a in
[0.45, 0.6] and b = c, ids permuted.  No reference counterpart: the
reference's generator has fixed probabilities and one seed
(gapbs/generator.h:89-186) and runs one graph per process.
"""
from __future__ import annotations

from typing import Iterator, Tuple

import numpy as np

from . import host as H

MANY_SEED = 20240


def many_small_graph_params(count: int, seed: int = MANY_SEED):
    """(scale, edge factor, a, b, c, generator seed) of each graph."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        scale = int(rng.integers(10, 17))
        ef = int(rng.choice([8, 16, 32]))
        a = float(rng.uniform(0.45, 0.6))
        b = c = float(rng.uniform(0.1, (1 - a) / 2 - 0.02))
        out.append((scale, ef, a, b, c, 5000 + i))
    return out


def many_small_graphs(count: int, seed: int = MANY_SEED) -> Iterator[Tuple[np.ndarray, np.ndarray, int]]:
    """Yields (u, v, n) edge lists (self loops dropped) of the batch."""
    for scale, ef, a, b, c, s in many_small_graph_params(count, seed):
        n = 1 << scale
        u, v = H.generate_rmat(scale, n * ef, a, b, c, s, True)
        keep = u != v
        yield u[keep], v[keep], n
