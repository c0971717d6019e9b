"""ctypes binding of oracle/_ref/libref.so: the reference's own header-only
code (generator, RelabelByMappingStandalone, mapping producers, RCM, the
GraphBrew pipeline, Tier-0 features) compiled from the reference tree by
oracle/Makefile (ref_glue.cc adds only C entry points).

TEST INFRASTRUCTURE ONLY.  Used by tests/ to pin the oracle and by
tests/golden/make_ref_vectors.py to write golden vectors and mapping
fixtures.  available() is False where the library was not built (the
reference tree absent); callers skip then.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libref.so")
REF_ROOT = "/root/reference"
_lib = None

I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
I64P = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
VP = C.c_void_p

GRAPHBREW_LEIDEN_GORDER8 = "leiden:compose:sg_none:comm_size_desc:intra_gorder:gw8"


def build() -> bool:
    """Compiles libref.so when the reference tree is present; True if built."""
    if not os.path.isdir(os.path.join(REF_ROOT, "bench", "include")):
        return False
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "_ref/libref.so", f"REF={REF_ROOT}"], check=True)
    return True


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        i32, i64 = C.c_int, C.c_int64
        sig = {
            "ref_generate_el": (i32, [i32, i32, i32, I32P, I32P]),
            "ref_insert_weights": (i32, [i64, I32P]),
            "ref_relabel": (i32, [i64, i32, I64P, I32P, VP, VP, VP, VP, I32P,
                                  I64P, I32P, VP, VP, VP, VP]),
            "ref_mapping": (i32, [i32, i64, i32, I64P, I32P, VP, VP, I32P]),
            "ref_graphbrew_mapping": (i32, [C.c_char_p, i64, i32, I64P, I32P, VP, VP, I32P]),
            "ref_tier0_features": (i32, [i64, i32, I64P, I32P, VP, VP, i64, C.c_uint64, F64P]),
            "ref_modeled_property_bytes": (i64, [i32, i64, i64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(VP)


def generate_el(scale: int, degree: int, uniform: bool = False):
    m = (1 << scale) * degree
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    lib().ref_generate_el(scale, degree, int(uniform), u, v)
    return u, v


def insert_weights(m: int) -> np.ndarray:
    w = np.empty(m, np.int32)
    lib().ref_insert_weights(m, w)
    return w


def _in(g):
    if g.directed:
        return _p(np.ascontiguousarray(g.in_off)), _p(np.ascontiguousarray(g.in_nbr))
    return None, None


def relabel(g, new_ids):
    """(off, nbr, wgt, in_off, in_nbr, in_wgt) of the reference's relabel."""
    n, m = g.n, int(g.off[-1])
    ids = np.ascontiguousarray(new_ids, np.int32)
    off = np.empty(n + 1, np.int64)
    nbr = np.empty(m, np.int32)
    w = np.empty(m, np.int32) if g.wgt is not None else None
    mi = int(g.in_off[-1]) if g.directed else 0
    in_off = np.empty(n + 1, np.int64) if g.directed else None
    in_nbr = np.empty(mi, np.int32) if g.directed else None
    in_w = np.empty(mi, np.int32) if (g.directed and g.wgt is not None) else None
    keep = [g.wgt, g.in_off, g.in_nbr, g.in_wgt]
    lib().ref_relabel(n, int(g.directed), g.off, g.nbr, _p(g.wgt),
                      _p(g.in_off) if g.directed else None, _p(g.in_nbr) if g.directed else None,
                      _p(g.in_wgt) if g.directed else None, ids, off, nbr, _p(w), _p(in_off),
                      _p(in_nbr), _p(in_w))
    del keep
    return off, nbr, w, in_off, in_nbr, in_w


def mapping(g, algo: int) -> np.ndarray:
    ids = np.empty(g.n, np.int32)
    io, inb = _in(g)
    rc = lib().ref_mapping(int(algo), g.n, int(g.directed), g.off, g.nbr, io, inb, ids)
    if rc != 0:
        raise ValueError(f"reference mapping {algo} unavailable")
    return ids


def graphbrew_mapping(g, options: str = GRAPHBREW_LEIDEN_GORDER8) -> np.ndarray:
    ids = np.empty(g.n, np.int32)
    io, inb = _in(g)
    rc = lib().ref_graphbrew_mapping(options.encode(), g.n, int(g.directed), g.off, g.nbr, io, inb, ids)
    if rc != 0:
        raise ValueError(f"graphbrew options {options!r} not on the delegated path")
    return ids


def tier0_features(g, sample_size: int = 0, seed: int = 0) -> np.ndarray:
    f = np.empty(7, np.float64)
    io, inb = _in(g)
    lib().ref_tier0_features(g.n, int(g.directed), g.off, g.nbr, io, inb, sample_size, seed, f)
    return f


def modeled_property_bytes(bench: int, n: int, m: int) -> int:
    return int(lib().ref_modeled_property_bytes(bench, n, m))
