"""ctypes binding of libgbkhost.so (include/gbk_host.h): the host substrate
of the bench path -- synthetic graphs, edge-list files, mapping producers and
files, source picking, the Tier-0 features and the frozen rule."""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libgbkhost.so")
_lib = None
VP = C.c_void_p

# BenchmarkType codes (reorder_types.h:1214-1224)
BENCH_CODES = {"pr": 1, "bfs": 2, "cc": 3, "sssp": 4, "bc": 5, "tc": 6, "pr_spmv": 7, "cc_sv": 8}
TIER0_NAMES = ("avg_degree", "degree_cv", "hub_concentration", "normalized_edge_span",
               "window_neighbor_overlap", "packing_factor", "internal_density")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run graphbrew_amd.build.build()")
        L = C.CDLL(LIB_PATH)
        i32, i64, f64 = C.c_int, C.c_int64, C.c_double
        sig = {
            "gbh_generate_el": (i32, [i32, i32, i32, VP, VP]),
            "gbh_insert_weights": (i32, [i64, VP]),
            "gbh_read_el": (i32, [C.c_char_p, i32, C.POINTER(i64), VP, VP, VP]),
            "gbh_mapping_identity": (i32, [i64, VP]),
            "gbh_mapping_random": (i32, [i64, VP]),
            "gbh_mapping_sort": (i32, [i64, VP, VP]),
            "gbh_mapping_dbg": (i32, [i64, VP, VP]),
            "gbh_mapping_hubsort": (i32, [i64, VP, VP]),
            "gbh_mapping_hubcluster": (i32, [i64, VP, VP]),
            "gbh_mapping_hubsortdbg": (i32, [i64, VP, VP]),
            "gbh_mapping_hubclusterdbg": (i32, [i64, VP, VP]),
            "gbh_mapping_fingerprint": (C.c_uint64, [i64, VP]),
            "gbh_generate_rmat": (i32, [i32, i64, f64, f64, f64, C.c_uint64, i32, VP, VP]),
            "gbh_verify_bfs": (i32, [i64, VP, VP, VP, VP, C.c_int32, VP]),
            "gbh_verify_cc": (i32, [i64, VP, VP, VP, VP, i32, VP]),
            "gbh_verify_sssp": (i64, [i64, VP, VP, VP, C.c_int32, VP]),
            "gbh_verify_bc": (i32, [i64, VP, VP, VP, i32, VP, C.POINTER(f64)]),
            "gbh_pr_residual": (f64, [i64, VP, VP, VP]),
            "gbh_weight_checksum": (None, [i64, VP, VP, VP, VP, VP]),
            "gbh_distance_fingerprint": (None, [i64, VP, VP, VP]),
            "gbh_validate_permutation": (i32, [i64, VP]),
            "gbh_mapping_read": (i32, [C.c_char_p, C.POINTER(i64), VP]),
            "gbh_sg_info": (i32, [C.c_char_p, C.POINTER(i32), C.POINTER(i64), C.POINTER(i64), C.POINTER(i32)]),
            "gbh_sg_read": (i32, [C.c_char_p, i32, VP, VP, VP, VP, VP, VP, VP]),
            "gbh_sg_write": (i32, [C.c_char_p, i32, i64, i64, VP, VP, VP, VP, VP, VP, VP]),
            "gbh_mapping_write": (i32, [C.c_char_p, i64, VP]),
            "gbh_source_picker": (i32, [i64, VP, VP, i32, i32, VP, VP]),
            "gbh_afforest_samples": (i32, [i64, i32, VP]),
            "gbh_tier0_features": (i32, [i64, VP, VP, i64, C.c_uint64, VP]),
            "gbh_modeled_property_bytes": (i64, [i32, i64, i64]),
            "gbh_frozen_rule": (i32, [i32, f64, i64, f64, f64, f64]),
            "gbh_budgeted_rule": (i32, [i32, f64, i64, f64, f64, f64]),
            "gbh_generate_lattice": (i64, [i64, i64, f64, C.c_uint64, i32, VP, VP, VP]),
            "gbh_llc_bytes": (i64, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(VP)


def _chk(rc, what):
    if rc < 0:
        raise ValueError(f"{what} failed with code {rc}")


def generate_el(scale: int, degree: int = 16, uniform: bool = False):
    m = (1 << scale) * degree
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    _chk(lib().gbh_generate_el(scale, degree, int(uniform), _p(u), _p(v)), "generate_el")
    return u, v


def generate_rmat(scale: int, edges: int, a: float, b: float, c: float, seed: int, permute: bool):
    u = np.empty(edges, np.int32)
    v = np.empty(edges, np.int32)
    _chk(lib().gbh_generate_rmat(scale, edges, a, b, c, seed, int(permute), _p(u), _p(v)), "generate_rmat")
    return u, v


def generate_lattice(rows: int, cols: int, rewire: float = 0.001, seed: int = 1, max_weight: int = 1000,
                     weighted: bool = True):
    """Road-like grid (4-neighbour, rewired heads, weights U[1, max_weight]):
    (u, v, w or None), n = rows * cols."""
    m = int(lib().gbh_generate_lattice(rows, cols, rewire, seed, max_weight, None, None, None))
    if m < 0:
        raise ValueError("bad lattice arguments")
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    w = np.empty(m, np.int32) if weighted else None
    _chk(lib().gbh_generate_lattice(rows, cols, rewire, seed, max_weight, _p(u), _p(v), _p(w)) - m, "lattice")
    return u, v, w


def insert_weights(m: int) -> np.ndarray:
    w = np.empty(m, np.int32)
    _chk(lib().gbh_insert_weights(m, _p(w)), "insert_weights")
    return w


def read_el(path: str, weighted: bool = False):
    m = C.c_int64(0)
    _chk(lib().gbh_read_el(path.encode(), int(weighted), C.byref(m), None, None, None), "read_el")
    u = np.empty(m.value, np.int32)
    v = np.empty(m.value, np.int32)
    w = np.empty(m.value, np.int32) if weighted else None
    _chk(lib().gbh_read_el(path.encode(), int(weighted), C.byref(m), _p(u), _p(v), _p(w)), "read_el")
    return u, v, w


def mapping_identity(n: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    _chk(lib().gbh_mapping_identity(n, _p(out)), "identity")
    return out


def mapping_random(n: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    _chk(lib().gbh_mapping_random(n, _p(out)), "random")
    return out


def mapping_sort(off: np.ndarray) -> np.ndarray:
    off = np.ascontiguousarray(off, np.int64)
    out = np.empty(len(off) - 1, np.int32)
    _chk(lib().gbh_mapping_sort(len(off) - 1, _p(off), _p(out)), "sort")
    return out


def mapping_dbg(off: np.ndarray) -> np.ndarray:
    off = np.ascontiguousarray(off, np.int64)
    out = np.empty(len(off) - 1, np.int32)
    _chk(lib().gbh_mapping_dbg(len(off) - 1, _p(off), _p(out)), "dbg")
    return out


def _hub_mapping(name, off):
    off = np.ascontiguousarray(off, np.int64)
    out = np.empty(len(off) - 1, np.int32)
    _chk(getattr(lib(), "gbh_mapping_" + name)(len(off) - 1, _p(off), _p(out)), name)
    return out


def mapping_hubsort(off: np.ndarray) -> np.ndarray:
    return _hub_mapping("hubsort", off)


def mapping_hubcluster(off: np.ndarray) -> np.ndarray:
    return _hub_mapping("hubcluster", off)


def mapping_hubsortdbg(off: np.ndarray) -> np.ndarray:
    return _hub_mapping("hubsortdbg", off)


def mapping_hubclusterdbg(off: np.ndarray) -> np.ndarray:
    return _hub_mapping("hubclusterdbg", off)


def mapping(algo: int, off: np.ndarray) -> np.ndarray:
    """new_ids of the cheap orderings -o 0..7 (original, random, sort,
    hubsort, hubcluster, dbg, hubsortdbg, hubclusterdbg)."""
    n = len(off) - 1
    if algo == 0:
        return mapping_identity(n)
    if algo == 1:
        return mapping_random(n)
    fn = {2: mapping_sort, 3: mapping_hubsort, 4: mapping_hubcluster, 5: mapping_dbg,
          6: mapping_hubsortdbg, 7: mapping_hubclusterdbg}.get(algo)
    if fn is None:
        raise ValueError(f"ordering {algo} is not produced here")
    return fn(off)


def mapping_fingerprint(ids) -> str:
    ids = np.ascontiguousarray(ids, np.int32)
    return f"{lib().gbh_mapping_fingerprint(len(ids), _p(ids)):016x}"


def validate_permutation(ids) -> int:
    ids = np.ascontiguousarray(ids, np.int32)
    return lib().gbh_validate_permutation(len(ids), _p(ids))


def mapping_read(path: str) -> np.ndarray:
    n = C.c_int64(0)
    _chk(lib().gbh_mapping_read(path.encode(), C.byref(n), None), "mapping_read")
    out = np.empty(n.value, np.int32)
    _chk(lib().gbh_mapping_read(path.encode(), C.byref(n), _p(out)), "mapping_read")
    return out


def mapping_write(path: str, ids) -> None:
    ids = np.ascontiguousarray(ids, np.int32)
    _chk(lib().gbh_mapping_write(path.encode(), len(ids), _p(ids)), "mapping_write")


def sg_info(path: str) -> dict:
    """Header of a .sg / .wsg serialized graph (gapbs/reader.h:445-485)."""
    d, w = C.c_int32(0), C.c_int32(0)
    n, m = C.c_int64(0), C.c_int64(0)
    _chk(lib().gbh_sg_info(path.encode(), C.byref(d), C.byref(n), C.byref(m), C.byref(w)), "sg_info")
    return {"directed": bool(d.value), "n": n.value, "m": m.value, "weighted": bool(w.value)}


def sg_read(path: str, weights: bool = False, scheme: str = "unit") -> dict:
    """Reads a serialized graph: off, nbr, org_ids (+ in_off, in_nbr for a
    directed graph; + wgt, in_wgt when weights, generated by 'scheme' for a
    .sg file as ReadSerializedGraph does, gapbs/reader.h:48-59)."""
    h = sg_info(path)
    n, m, d = h["n"], h["m"], h["directed"]
    out = {"directed": d, "off": np.zeros(n + 1, np.int64), "nbr": np.zeros(m, np.int32),
           "org_ids": np.zeros(n, np.int32)}
    if d:
        out["in_off"] = np.zeros(n + 1, np.int64)
        out["in_nbr"] = np.zeros(m, np.int32)
    if weights:
        out["wgt"] = np.zeros(m, np.int32)
        if d:
            out["in_wgt"] = np.zeros(m, np.int32)
    _chk(lib().gbh_sg_read(path.encode(), {"unit": 0, "hash": 1}[scheme], _p(out["off"]), _p(out["nbr"]),
                           _p(out.get("wgt")), _p(out.get("in_off")), _p(out.get("in_nbr")),
                           _p(out.get("in_wgt")), _p(out["org_ids"])), "sg_read")
    return out


def sg_write(path: str, off, nbr, wgt=None, in_off=None, in_nbr=None, in_wgt=None, org_ids=None) -> None:
    """Writes a .sg (wgt None) or .wsg serialized graph (gapbs/writer.h:336-377);
    a directed graph passes its in-CSR."""
    off = np.ascontiguousarray(off, np.int64)
    nbr = np.ascontiguousarray(nbr, np.int32)
    n, m = len(off) - 1, len(nbr)
    org = np.arange(n, dtype=np.int32) if org_ids is None else np.ascontiguousarray(org_ids, np.int32)
    arr = lambda x, t: None if x is None else np.ascontiguousarray(x, t)
    wgt, in_off, in_nbr, in_wgt = arr(wgt, np.int32), arr(in_off, np.int64), arr(in_nbr, np.int32), arr(in_wgt, np.int32)
    _chk(lib().gbh_sg_write(path.encode(), int(in_off is not None), n, m, _p(off), _p(nbr), _p(wgt), _p(in_off),
                            _p(in_nbr), _p(in_wgt), _p(org)), "sg_write")


def source_picker(off, count: int, org_ids=None, repeats: int = 1):
    off = np.ascontiguousarray(off, np.int64)
    org = None if org_ids is None else np.ascontiguousarray(org_ids, np.int32)
    o = np.empty(count, np.int32)
    i = np.empty(count, np.int32)
    _chk(lib().gbh_source_picker(len(off) - 1, _p(off), _p(org), count, repeats, _p(o), _p(i)),
         "source_picker")
    return o, i


def afforest_samples(n: int, count: int = 1024) -> np.ndarray:
    out = np.empty(count, np.int32)
    _chk(lib().gbh_afforest_samples(n, count, _p(out)), "afforest_samples")
    return out


def tier0_features(off, nbr, sample_size=0, seed=0) -> dict:
    off = np.ascontiguousarray(off, np.int64)
    nbr = np.ascontiguousarray(nbr, np.int32)
    f = np.zeros(7, np.float64)
    _chk(lib().gbh_tier0_features(len(off) - 1, _p(off), _p(nbr), sample_size, seed, _p(f)), "tier0")
    return dict(zip(TIER0_NAMES, f.tolist()))


def modeled_property_bytes(bench: str, n: int, m: int) -> int:
    return int(lib().gbh_modeled_property_bytes(BENCH_CODES[bench], n, m))


def frozen_rule(bench: str, reuse: float, n: int, avg_degree: float, degree_cv: float,
                wsr: float) -> bool:
    return bool(lib().gbh_frozen_rule(BENCH_CODES[bench], reuse, n, avg_degree, degree_cv, wsr))


def budgeted_rule(bench: str, reuse: float, n: int, avg_degree: float, degree_cv: float,
                  hub_concentration: float) -> bool:
    return bool(lib().gbh_budgeted_rule(BENCH_CODES[bench], reuse, n, avg_degree, degree_cv, hub_concentration))


def llc_bytes() -> int:
    return int(lib().gbh_llc_bytes())


def rule_pick(rule: str, bench: str, reuse: float, n: int, m: int, tier0: dict, llc: int = 0):
    """The reference's deployable rules (reorder_adaptive.h:240-383):
    "allkernel-lowreuse-rule" asks for the Leiden-Gorder8 composition when
    its predicate holds, "budgeted-rule" for the budgeted Leiden-Gorder
    composition; otherwise both fall back to DBG (-o 5), their arm in builds
    without Rabbit order.  Returns (wants_leiden, working-set ratio)."""
    llc = llc or llc_bytes()
    wsr = modeled_property_bytes(bench, n, m) / llc if llc > 0 else 0.0
    if rule == "allkernel-lowreuse-rule":
        hit = frozen_rule(bench, reuse, n, tier0["avg_degree"], tier0["degree_cv"], wsr)
    elif rule == "budgeted-rule":
        hit = budgeted_rule(bench, reuse, n, tier0["avg_degree"], tier0["degree_cv"], tier0["hub_concentration"])
    else:
        raise ValueError(rule)
    return hit, wsr


# ------------------------------------------------------------- verifiers
def _a32(x):
    return None if x is None else np.ascontiguousarray(x, np.int32)


def verify_bfs(off, nbr, in_off, in_nbr, source: int, parent) -> bool:
    return bool(lib().gbh_verify_bfs(len(off) - 1, _p(off), _p(nbr), _p(in_off), _p(in_nbr), source,
                                     _p(_a32(parent))))


def verify_cc(off, nbr, in_off, in_nbr, directed: bool, comp) -> bool:
    return bool(lib().gbh_verify_cc(len(off) - 1, _p(off), _p(nbr), _p(in_off), _p(in_nbr),
                                    int(directed), _p(_a32(comp))))


def verify_sssp(off, nbr, wgt, source: int, dist) -> int:
    return int(lib().gbh_verify_sssp(len(off) - 1, _p(off), _p(nbr), _p(wgt), source, _p(_a32(dist))))


def verify_bc(off, nbr, sources, scores):
    s = _a32(sources)
    sc = np.ascontiguousarray(scores, np.float32)
    worst = C.c_double(0)
    ok = lib().gbh_verify_bc(len(off) - 1, _p(off), _p(nbr), _p(s), len(s), _p(sc), C.byref(worst))
    return bool(ok), worst.value


def pr_residual(off, nbr, scores) -> float:
    return float(lib().gbh_pr_residual(len(off) - 1, _p(off), _p(nbr),
                                       _p(np.ascontiguousarray(scores, np.float32))))


def weight_checksum(off, nbr, wgt, org_ids=None) -> str:
    out = np.zeros(2, np.uint64)
    lib().gbh_weight_checksum(len(off) - 1, _p(off), _p(nbr), _p(wgt), _p(_a32(org_ids)), _p(out))
    return f"{int(out[0]):016x}{int(out[1]):016x}"


def distance_fingerprint(dist, org_ids=None) -> str:
    d = _a32(dist)
    out = np.zeros(2, np.uint64)
    lib().gbh_distance_fingerprint(len(d), _p(d), _p(_a32(org_ids)), _p(out))
    return f"{int(out[0]):016x}{int(out[1]):016x}"
