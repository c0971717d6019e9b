"""Independent textbook algorithms used to pin the oracle on small inputs.

Written from the mathematics (set-based CSR, BFS, Floyd-Warshall, union-find,
Brandes' dependency recursion, a dense linear solve for PageRank) rather than
from the reference's code, so that a mistake in the oracle's restatement
(dropped term, wrong index, transposed operand) shows up as a mismatch.
"""
from __future__ import annotations

import collections
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

MT_A, MT_N, MT_M = 0x9908B0DF, 624, 397


class MT19937:
    """MT19937 from Matsumoto & Nishimura (1998), the generator std::mt19937 names."""

    def __init__(self, seed: int = 5489):
        self.mt = [0] * MT_N
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, MT_N):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.idx = MT_N

    def _twist(self):
        mt = self.mt
        for i in range(MT_N):
            y = (mt[i] & 0x80000000) | (mt[(i + 1) % MT_N] & 0x7FFFFFFF)
            mt[i] = mt[(i + MT_M) % MT_N] ^ (y >> 1) ^ (MT_A if y & 1 else 0)
        self.idx = 0

    def __call__(self) -> int:
        if self.idx >= MT_N:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def splitmix64_stream(state: int, count: int) -> List[int]:
    """SplitMix64 (Steele, Lea, Flood 2014)."""
    out = []
    M = (1 << 64) - 1
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & M
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        out.append(z ^ (z >> 31))
    return out


def read_el(path: str, weighted: bool = False):
    us, vs, ws = [], [], []
    with open(path) as f:
        for line in f:
            if not line.strip() or line[0] in "%#":
                continue
            parts = line.split()
            us.append(int(parts[0]))
            vs.append(int(parts[1]))
            if weighted:
                ws.append(int(parts[2]))
    u = np.array(us, np.int32)
    v = np.array(vs, np.int32)
    return (u, v, np.array(ws, np.int32)) if weighted else (u, v, None)


def adjacency_sets(n, u, v, w=None, symmetrize=True, keep_self=False):
    """Row -> {neighbour: smallest weight}; the squished CSR as sets."""
    rows: List[Dict[int, int]] = [dict() for _ in range(n)]
    for i in range(len(u)):
        a, b = int(u[i]), int(v[i])
        wt = 1 if w is None else int(w[i])
        pairs = [(a, b), (b, a)] if symmetrize else [(a, b)]
        for x, y in pairs:
            if x == y and not keep_self:
                continue
            if y not in rows[x] or wt < rows[x][y]:
                rows[x][y] = wt
    return rows


def csr_from_sets(rows):
    off = [0]
    nbr, wgt = [], []
    for r in rows:
        for k in sorted(r):
            nbr.append(k)
            wgt.append(r[k])
        off.append(len(nbr))
    return np.array(off, np.int64), np.array(nbr, np.int32), np.array(wgt, np.int32)


def neighbors(off, nbr, u):
    return nbr[off[u]:off[u + 1]]


def bfs_depths(n, off, nbr, s):
    d = [-1] * n
    d[s] = 0
    q = collections.deque([s])
    while q:
        x = q.popleft()
        for y in neighbors(off, nbr, x):
            if d[y] < 0:
                d[y] = d[x] + 1
                q.append(y)
    return np.array(d, np.int32)


def floyd_warshall(n, off, nbr, wgt):
    INF = np.int64(1) << 40
    D = np.full((n, n), INF, np.int64)
    np.fill_diagonal(D, 0)
    for x in range(n):
        for k in range(off[x], off[x + 1]):
            D[x, nbr[k]] = min(D[x, nbr[k]], wgt[k])
    for k in range(n):
        D = np.minimum(D, D[:, k:k + 1] + D[k:k + 1, :])
    return D, INF


def components_min_label(n, off, nbr, in_off=None, in_nbr=None):
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for x in range(n):
        for y in neighbors(off, nbr, x):
            a, b = find(x), find(int(y))
            if a != b:
                parent[max(a, b)] = min(a, b)
    lab = np.array([find(x) for x in range(n)], np.int32)
    # label = smallest vertex id in the component
    mins = {}
    for x in range(n):
        mins.setdefault(lab[x], x)
    return np.array([mins[lab[x]] for x in range(n)], np.int32)


def brandes(n, off, nbr, sources: Sequence[int]) -> np.ndarray:
    """Brandes (2001) dependency accumulation, exact rationals via float64."""
    bc = np.zeros(n, np.float64)
    for s in sources:
        dist = [-1] * n
        sigma = [0] * n
        dist[s], sigma[s] = 0, 1
        order = []
        q = collections.deque([s])
        while q:
            x = q.popleft()
            order.append(x)
            for y in neighbors(off, nbr, x):
                y = int(y)
                if dist[y] < 0:
                    dist[y] = dist[x] + 1
                    q.append(y)
                if dist[y] == dist[x] + 1:
                    sigma[y] += sigma[x]
        delta = [0.0] * n
        for x in reversed(order):
            for y in neighbors(off, nbr, x):
                y = int(y)
                if dist[y] == dist[x] + 1:
                    delta[x] += sigma[x] / sigma[y] * (1.0 + delta[y])
            bc[x] += delta[x]
    mx = bc.max() if n else 0.0
    return bc / mx if mx > 0 else bc


def pagerank_dense(n, off, nbr, d=0.85):
    """Solve x = (1-d)/n + d * A^T D^-1 x exactly (dangling mass leaks, as in GAPBS)."""
    M = np.zeros((n, n))
    for x in range(n):
        deg = off[x + 1] - off[x]
        for k in range(off[x], off[x + 1]):
            M[nbr[k], x] += 1.0 / deg
    return np.linalg.solve(np.eye(n) - d * M, np.full(n, (1 - d) / n))


def random_graph(n: int, m: int, seed: int, weighted=False, wmax=255):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m).astype(np.int32)
    v = rng.integers(0, n, m).astype(np.int32)
    w = rng.integers(1, wmax + 1, m).astype(np.int32) if weighted else None
    return u, v, w
