"""ctypes binding of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg as the timed CPU baseline; the
product (graphbrew_amd/) never imports it.

Each wrapper takes and returns numpy arrays; CSR offsets are int64, ids and
weights int32.  The C functions cite the reference file:line they restate.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
I64P = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
U32P = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
U64P = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
VP = C.c_void_p


def build() -> str:
    """Compile liboracle.so from oracle.cpp (make, in this directory)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        i64, i32, f64 = C.c_int64, C.c_int, C.c_double
        sig = {
            "orc_generate_el": (i32, [i32, i32, i32, I32P, I32P]),
            "orc_insert_weights": (i32, [i64, I32P]),
            "orc_mt19937": (None, [C.c_uint32, i64, U32P]),
            "orc_max_node_id": (C.c_int32, [i64, I32P, I32P]),
            "orc_build_csr": (i64, [i64, i64, I32P, I32P, VP, i32, i32, i32, I64P, I32P, VP]),
            "orc_relabel": (i32, [i64, I64P, I32P, VP, I32P, I64P, I32P, VP]),
            "orc_update_org_ids": (i32, [i64, VP, I32P, I32P]),
            "orc_mapping_fingerprint": (C.c_uint64, [i64, I32P]),
            "orc_validate_permutation": (i32, [i64, I32P]),
            "orc_mapping_random": (i32, [i64, I32P]),
            "orc_mapping_sort": (i32, [i64, I64P, I32P]),
            "orc_mapping_dbg": (i32, [i64, I64P, I32P]),
            "orc_source_picker": (i32, [i64, I64P, VP, i32, i32, I32P, I32P]),
            "orc_pr_gs": (i32, [i64, I64P, I32P, I64P, i32, f64, i32, i32, F32P,
                                C.POINTER(C.c_int), C.POINTER(C.c_double)]),
            "orc_pr_spmv": (i32, [i64, I64P, I32P, I64P, i32, f64, i32, i32, F32P,
                                  C.POINTER(C.c_int), C.POINTER(C.c_double)]),
            "orc_pr_residual": (f64, [i64, I64P, I32P, F32P]),
            "orc_pr_fp64": (i32, [i64, I64P, I32P, I64P, i32, f64, F64P]),
            "orc_bfs_do": (i32, [i64, I64P, I32P, I64P, I32P, C.c_int32, i32, i32, i32, I32P]),
            "orc_bfs_depths": (i32, [i64, I64P, I32P, C.c_int32, I32P]),
            "orc_bfs_verify": (i32, [i64, I64P, I32P, I64P, I32P, C.c_int32, I32P]),
            "orc_cc_afforest": (i32, [i64, I64P, I32P, I64P, I32P, i32, i32, i32, I32P]),
            "orc_cc_sv": (i32, [i64, I64P, I32P, i32, I32P]),
            "orc_cc_verify": (i32, [i64, I64P, I32P, I64P, I32P, i32, I32P]),
            "orc_afforest_sample_indices": (i32, [i64, i32, I32P]),
            "orc_sssp_delta": (i32, [i64, I64P, I32P, I32P, C.c_int32, C.c_int32, i32, I32P]),
            "orc_dijkstra": (i32, [i64, I64P, I32P, I32P, C.c_int32, I32P]),
            "orc_weight_checksum": (None, [i64, I64P, I32P, I32P, VP, U64P]),
            "orc_distance_fingerprint": (None, [i64, I32P, VP, U64P]),
            "orc_bc": (i32, [i64, I64P, I32P, I32P, i32, i32, F32P]),
            "orc_bc_serial": (i32, [i64, I64P, I32P, I32P, i32, F32P]),
            "orc_bc_fp64": (i32, [i64, I64P, I32P, I32P, i32, i32, F64P]),
            "orc_bc_within_tolerance": (i32, [C.c_float, C.c_float]),
            "orc_tier0_features": (i32, [i64, i64, I64P, I32P, i64, C.c_uint64, F64P]),
            "orc_modeled_property_bytes": (i64, [i32, i64, i64]),
            "orc_frozen_rule": (i32, [i32, f64, i64, f64, f64, f64]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _vp(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc: int, what: str):
    if rc < 0:
        raise ValueError(f"{what} failed with code {rc}")


@dataclass
class Graph:
    """Host CSR (GAPBS CSRGraph layout: int64 offsets, int32 neighbours)."""
    n: int
    off: np.ndarray
    nbr: np.ndarray
    wgt: Optional[np.ndarray] = None
    in_off: Optional[np.ndarray] = None
    in_nbr: Optional[np.ndarray] = None
    in_wgt: Optional[np.ndarray] = None
    directed: bool = False
    org_ids: Optional[np.ndarray] = None

    @property
    def m(self) -> int:
        """Directed edge count (num_edges_directed, gapbs/graph.h:377-380)."""
        return int(self.off[-1] - self.off[0])

    def inv(self):
        """(in_off, in_nbr): the out arrays for an undirected graph."""
        if self.directed:
            return self.in_off, self.in_nbr
        return self.off, self.nbr


# ---------------------------------------------------------------- generator
def generate_el(scale: int, degree: int, uniform: bool = False):
    m = (1 << scale) * degree
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    _check(lib().orc_generate_el(scale, degree, int(uniform), u, v), "generate_el")
    return u, v


def insert_weights(m: int) -> np.ndarray:
    w = np.empty(m, np.int32)
    _check(lib().orc_insert_weights(m, w), "insert_weights")
    return w


def mt19937(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, np.uint32)
    lib().orc_mt19937(seed, count, out)
    return out


def build_csr(n: int, u, v, w=None, symmetrize=True, transpose=False, keep_self=False):
    u = np.ascontiguousarray(u, np.int32)
    v = np.ascontiguousarray(v, np.int32)
    m = len(u)
    cap = 2 * m if symmetrize else m
    off = np.empty(n + 1, np.int64)
    nbr = np.empty(max(cap, 1), np.int32)
    wgt = np.empty(max(cap, 1), np.int32) if w is not None else None
    wv = None if w is None else np.ascontiguousarray(w, np.int32)
    cnt = lib().orc_build_csr(n, m, u, v, _vp(wv), int(symmetrize), int(transpose),
                              int(keep_self), off, nbr, _vp(wgt))
    _check(cnt, "build_csr")
    return off, nbr[:cnt].copy(), (wgt[:cnt].copy() if wgt is not None else None)


def make_graph(u, v, w=None, symmetrize=True, keep_self=False, n=None) -> Graph:
    """MakeGraphFromEL + SquishGraph (gapbs/builder.h:463-493, 292-318)."""
    u = np.ascontiguousarray(u, np.int32)
    v = np.ascontiguousarray(v, np.int32)
    if n is None:
        n = int(lib().orc_max_node_id(len(u), u, v)) + 1
    off, nbr, wgt = build_csr(n, u, v, w, symmetrize, False, keep_self)
    g = Graph(n=n, off=off, nbr=nbr, wgt=wgt, directed=not symmetrize)
    if not symmetrize:
        g.in_off, g.in_nbr, g.in_wgt = build_csr(n, u, v, w, False, True, keep_self)
    return g


def kron(scale: int, degree: int = 16, uniform: bool = False, weighted: bool = False) -> Graph:
    """The -g/-u generated graph: symmetrized, squished (command_line.h:114-115)."""
    u, v = generate_el(scale, degree, uniform)
    w = insert_weights(len(u)) if weighted else None
    return make_graph(u, v, w, symmetrize=True)


# ------------------------------------------------------------------ relabel
def relabel_csr(n, off, nbr, wgt, new_ids):
    new_ids = np.ascontiguousarray(new_ids, np.int32)
    off_o = np.empty(n + 1, np.int64)
    nbr_o = np.empty(max(len(nbr), 1), np.int32)
    wgt_o = np.empty(max(len(nbr), 1), np.int32) if wgt is not None else None
    _check(lib().orc_relabel(n, off, nbr, _vp(wgt), new_ids, off_o, nbr_o, _vp(wgt_o)), "relabel")
    m = len(nbr)
    return off_o, nbr_o[:m].copy(), (wgt_o[:m].copy() if wgt is not None else None)


def relabel(g: Graph, new_ids) -> Graph:
    """RelabelByMapping (gapbs/builder.h:1046-1206) incl. org_ids update."""
    off, nbr, wgt = relabel_csr(g.n, g.off, g.nbr, g.wgt, new_ids)
    r = Graph(n=g.n, off=off, nbr=nbr, wgt=wgt, directed=g.directed)
    if g.directed:
        r.in_off, r.in_nbr, r.in_wgt = relabel_csr(g.n, g.in_off, g.in_nbr, g.in_wgt, new_ids)
    org = np.empty(g.n, np.int32)
    _check(lib().orc_update_org_ids(g.n, _vp(g.org_ids), np.ascontiguousarray(new_ids, np.int32), org),
           "update_org_ids")
    r.org_ids = org
    return r


def mapping_fingerprint(new_ids) -> str:
    h = lib().orc_mapping_fingerprint(len(new_ids), np.ascontiguousarray(new_ids, np.int32))
    return f"{h:016x}"


def validate_permutation(ids) -> int:
    return lib().orc_validate_permutation(len(ids), np.ascontiguousarray(ids, np.int32))


def mapping_random(n: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    _check(lib().orc_mapping_random(n, out), "mapping_random")
    return out


def mapping_sort(g: Graph) -> np.ndarray:
    out = np.empty(g.n, np.int32)
    _check(lib().orc_mapping_sort(g.n, g.off, out), "mapping_sort")
    return out


def mapping_dbg(g: Graph) -> np.ndarray:
    out = np.empty(g.n, np.int32)
    _check(lib().orc_mapping_dbg(g.n, g.off, out), "mapping_dbg")
    return out


def _hubs(off: np.ndarray):
    """Vertices with degree above the integer average (reorder_hub.h:141-149)."""
    n = len(off) - 1
    deg = np.diff(off)
    avg = (int(off[-1]) - int(off[0])) // n if n else 0
    return np.nonzero(deg > avg)[0].astype(np.int64), deg


def _assign_hubs(n: int, hubs) -> np.ndarray:
    """AssignHubOrderPreservingSourceIds, reorder_hub.h:36-86, step by step."""
    new = np.full(n, -1, np.int64)
    is_hub = np.zeros(n, bool)
    for pos, v in enumerate(hubs):
        is_hub[v] = True
        new[v] = pos
    deferred = []
    for old in range(len(hubs), n):
        if not is_hub[old]:
            new[old] = old
            continue
        p = new[old]
        if not is_hub[p] and new[p] == -1:
            new[p] = old
        else:
            deferred.append(old)
    it = iter(deferred)
    for old in range(len(hubs)):
        if new[old] == -1:
            new[old] = next(it)
    return new.astype(np.int32)


def _hubs_by_degree(off):
    hubs, deg = _hubs(off)
    # std::greater on (degree, id) pairs: degree descending, then id descending
    return sorted(hubs.tolist(), key=lambda v: (-int(deg[v]), -v))


def mapping_hubsort(off) -> np.ndarray:
    """GenerateHubSortMapping, reorder_hub.h:125-199."""
    off = np.asarray(off, np.int64)
    return _assign_hubs(len(off) - 1, _hubs_by_degree(off))


def mapping_hubcluster(off) -> np.ndarray:
    """GenerateHubClusterMapping, reorder_hub.h:203-245 (hubs in id order)."""
    off = np.asarray(off, np.int64)
    return _assign_hubs(len(off) - 1, _hubs(off)[0].tolist())


def _hubs_then_rest(n, hubs):
    new = np.full(n, -1, np.int64)
    new[np.asarray(hubs, np.int64)] = np.arange(len(hubs))
    rest = np.nonzero(new < 0)[0]
    new[rest] = len(hubs) + np.arange(len(rest))
    return new.astype(np.int32)


def mapping_hubsortdbg(off) -> np.ndarray:
    """GenerateHubSortDBGMapping, reorder_hub.h:380-502: sorted hubs first,
    then the other vertices in id order (the per-thread slices are
    contiguous and numbered in thread order)."""
    off = np.asarray(off, np.int64)
    return _hubs_then_rest(len(off) - 1, _hubs_by_degree(off))


def mapping_hubclusterdbg(off) -> np.ndarray:
    """GenerateHubClusterDBGMapping, reorder_hub.h:505-582: two buckets
    (degree <= avg, above), the upper first, each in id order."""
    off = np.asarray(off, np.int64)
    return _hubs_then_rest(len(off) - 1, _hubs(off)[0].tolist())


def source_picker(g: Graph, count: int, repeats: int = 1):
    o = np.empty(count, np.int32)
    i = np.empty(count, np.int32)
    _check(lib().orc_source_picker(g.n, g.off, _vp(g.org_ids), count, repeats, o, i), "source_picker")
    return o, i


# ---------------------------------------------------------------- PageRank
def _pr(fn, g: Graph, max_iters, tol, fixed, threads):
    in_off, in_nbr = g.inv()
    s = np.empty(g.n, np.float32)
    it = C.c_int(0)
    err = C.c_double(0)
    _check(fn(g.n, in_off, in_nbr, g.off, max_iters, tol, int(fixed), threads, s,
              C.byref(it), C.byref(err)), "pagerank")
    return s, it.value, err.value


def pr_gs(g: Graph, max_iters=20, tol=1e-4, fixed=False, threads=1):
    return _pr(lib().orc_pr_gs, g, max_iters, tol, fixed, threads)


def pr_spmv(g: Graph, max_iters=20, tol=1e-4, fixed=False, threads=1):
    return _pr(lib().orc_pr_spmv, g, max_iters, tol, fixed, threads)


def pr_residual(g: Graph, scores) -> float:
    return lib().orc_pr_residual(g.n, g.off, g.nbr, np.ascontiguousarray(scores, np.float32))


def pr_fp64(g: Graph, max_iters=1000, tol=1e-13) -> np.ndarray:
    in_off, in_nbr = g.inv()
    s = np.empty(g.n, np.float64)
    _check(lib().orc_pr_fp64(g.n, in_off, in_nbr, g.off, max_iters, tol, s), "pr_fp64")
    return s


# --------------------------------------------------------------------- BFS
def bfs_do(g: Graph, source: int, alpha=15, beta=18, threads=1) -> np.ndarray:
    in_off, in_nbr = g.inv()
    p = np.empty(g.n, np.int32)
    _check(lib().orc_bfs_do(g.n, g.off, g.nbr, in_off, in_nbr, source, alpha, beta, threads, p), "bfs")
    return p


def bfs_depths(g: Graph, source: int) -> np.ndarray:
    d = np.empty(g.n, np.int32)
    _check(lib().orc_bfs_depths(g.n, g.off, g.nbr, source, d), "bfs_depths")
    return d


def bfs_verify(g: Graph, source: int, parent) -> bool:
    in_off, in_nbr = g.inv()
    return bool(lib().orc_bfs_verify(g.n, g.off, g.nbr, in_off, in_nbr, source,
                                     np.ascontiguousarray(parent, np.int32)))


# ---------------------------------------------------------------------- CC
def cc_afforest(g: Graph, rounds=2, threads=1) -> np.ndarray:
    in_off, in_nbr = g.inv()
    c = np.empty(g.n, np.int32)
    _check(lib().orc_cc_afforest(g.n, g.off, g.nbr, in_off, in_nbr, int(g.directed), rounds,
                                 threads, c), "afforest")
    return c


def cc_sv(g: Graph, threads=1) -> np.ndarray:
    c = np.empty(g.n, np.int32)
    _check(lib().orc_cc_sv(g.n, g.off, g.nbr, threads, c), "sv")
    return c


def cc_verify(g: Graph, comp) -> bool:
    in_off, in_nbr = g.inv()
    return bool(lib().orc_cc_verify(g.n, g.off, g.nbr, in_off, in_nbr, int(g.directed),
                                    np.ascontiguousarray(comp, np.int32)))


def afforest_sample_indices(n: int, k: int = 1024) -> np.ndarray:
    out = np.empty(k, np.int32)
    _check(lib().orc_afforest_sample_indices(n, k, out), "sample")
    return out


# -------------------------------------------------------------------- SSSP
K_DIST_INF = (2 ** 31 - 1) // 2


def sssp_delta(g: Graph, source: int, delta: int = 1, threads=1) -> np.ndarray:
    d = np.empty(g.n, np.int32)
    _check(lib().orc_sssp_delta(g.n, g.off, g.nbr, g.wgt, source, delta, threads, d), "sssp")
    return d


def dijkstra(g: Graph, source: int) -> np.ndarray:
    d = np.empty(g.n, np.int32)
    _check(lib().orc_dijkstra(g.n, g.off, g.nbr, g.wgt, source, d), "dijkstra")
    return d


def weight_checksum(g: Graph) -> str:
    out = np.zeros(2, np.uint64)
    lib().orc_weight_checksum(g.n, g.off, g.nbr, g.wgt, _vp(g.org_ids), out)
    return f"{int(out[0]):016x}{int(out[1]):016x}"


def distance_fingerprint(g: Graph, dist) -> str:
    out = np.zeros(2, np.uint64)
    lib().orc_distance_fingerprint(g.n, np.ascontiguousarray(dist, np.int32), _vp(g.org_ids), out)
    return f"{int(out[0]):016x}{int(out[1]):016x}"


# ---------------------------------------------------------------------- BC
def bc(g: Graph, sources, threads=1) -> np.ndarray:
    s = np.ascontiguousarray(sources, np.int32)
    out = np.empty(g.n, np.float32)
    _check(lib().orc_bc(g.n, g.off, g.nbr, s, len(s), threads, out), "bc")
    return out


def bc_fp64(g: Graph, sources, threads=1) -> np.ndarray:
    """Brandes with double deltas and scores: the exact yardstick."""
    s = np.ascontiguousarray(sources, np.int32)
    out = np.empty(g.n, np.float64)
    _check(lib().orc_bc_fp64(g.n, g.off, g.nbr, s, len(s), threads, out), "bc_fp64")
    return out


def bc_serial(g: Graph, sources) -> np.ndarray:
    s = np.ascontiguousarray(sources, np.int32)
    out = np.empty(g.n, np.float32)
    _check(lib().orc_bc_serial(g.n, g.off, g.nbr, s, len(s), out), "bc_serial")
    return out


def bc_within_tolerance(expected: float, actual: float) -> bool:
    return bool(lib().orc_bc_within_tolerance(expected, actual))


# ------------------------------------------------------------------ Tier-0
TIER0_NAMES = ("avg_degree", "degree_cv", "hub_concentration", "normalized_edge_span",
               "window_neighbor_overlap", "packing_factor", "internal_density")


def tier0_features(g: Graph, sample_size=0, seed=0) -> dict:
    f = np.zeros(7, np.float64)
    _check(lib().orc_tier0_features(g.n, g.m, g.off, g.nbr, sample_size, seed, f), "tier0")
    return dict(zip(TIER0_NAMES, f.tolist()))


def modeled_property_bytes(bench: int, n: int, m: int) -> int:
    return int(lib().orc_modeled_property_bytes(bench, n, m))


def frozen_rule(bench: int, reuse: float, n: int, avg_degree: float, degree_cv: float,
                wsr: float) -> bool:
    return bool(lib().orc_frozen_rule(bench, reuse, n, avg_degree, degree_cv, wsr))


# ---- selector MLP (test infrastructure for the device kernels) -----------
HIDDEN = 64


def mlp_unpack(p: np.ndarray, d: int):
    """Flat parameters W1[64][d], b1, W2[64][64], b2, w3, b3 as float64."""
    p = np.asarray(p, np.float64)
    i = 0
    W1 = p[i:i + HIDDEN * d].reshape(HIDDEN, d); i += HIDDEN * d
    b1 = p[i:i + HIDDEN]; i += HIDDEN
    W2 = p[i:i + HIDDEN * HIDDEN].reshape(HIDDEN, HIDDEN); i += HIDDEN * HIDDEN
    b2 = p[i:i + HIDDEN]; i += HIDDEN
    w3 = p[i:i + HIDDEN]; i += HIDDEN
    b3 = p[i]
    return W1, b1, W2, b2, w3, b3


def mlp_forward(p, d, X):
    W1, b1, W2, b2, w3, b3 = mlp_unpack(p, d)
    a1 = np.maximum(np.asarray(X, np.float64) @ W1.T + b1, 0)
    a2 = np.maximum(a1 @ W2.T + b2, 0)
    return a2 @ w3 + b3


def mlp_grad(p, d, X, y):
    """Mean squared error and its gradient, by the chain rule in float64."""
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    W1, b1, W2, b2, w3, b3 = mlp_unpack(p, d)
    z1 = X @ W1.T + b1
    a1 = np.maximum(z1, 0)
    z2 = a1 @ W2.T + b2
    a2 = np.maximum(z2, 0)
    yh = a2 @ w3 + b3
    B = len(X)
    e = yh - y
    gy = 2.0 * e / B
    gw3 = a2.T @ gy
    gb3 = gy.sum()
    dz2 = np.outer(gy, w3) * (z2 > 0)
    gW2 = dz2.T @ a1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    gW1 = dz1.T @ X
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return float((e * e).mean()), g