"""ctypes binding of libgbk.so (include/gbk.h).

The library is the only compute path: there is no CPU fallback.  Importing
this module does not load it; the first Context() does, and raises
GbkUnavailable when libgbk.so is missing or no GPU is visible.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libgbk.so")

GBK_OK, GBK_EINVAL, GBK_EPERM_MAP, GBK_EHIP, GBK_ENOMEM, GBK_ESTATE = 0, -1, -2, -3, -4, -5
PR_GS, PR_SPMV = 0, 1
CC_AFFOREST, CC_SV = 0, 1
KT_NONE, KT_PR_ITER, KT_SSSP_RELAX, KT_BFS_STEP, KT_BC_LEVEL = 0, 1, 2, 3, 4
DIST_INF = (2 ** 31 - 1) // 2


class GbkUnavailable(RuntimeError):
    pass


class GbkError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gbk error {code}: {msg}")
        self.code = code


class PrStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("iterations", C.c_int), ("converged_iter", C.c_int),
                ("final_error", C.c_double)]


class BfsStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("steps_td", C.c_int), ("steps_bu", C.c_int),
                ("reached", C.c_int64), ("edges_reached", C.c_int64)]


class CcStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("iterations", C.c_int), ("skipped_label", C.c_int32)]


class SsspStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("buckets", C.c_int64), ("relax_rounds", C.c_int64),
                ("reached", C.c_int64), ("edges_reached", C.c_int64)]


class BcStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("edges_reached", C.c_int64), ("max_depth", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("rounds", C.c_int64), ("buckets", C.c_int64),
                ("edges_reached", C.c_int64), ("rows_scanned", C.c_int64), ("edges_scanned", C.c_int64)]


class RelabelStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("n", C.c_int64), ("m", C.c_int64)]


_lib = None
VP = C.c_void_p


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(VP)


def lib():
    """Loads libgbk.so and declares its signatures (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GbkUnavailable(f"{LIB_PATH} is not built; run graphbrew_amd.build.build()")
    L = C.CDLL(LIB_PATH)
    i32, i64, f64, pp = C.c_int, C.c_int64, C.c_double, C.POINTER(VP)
    sig = {
        "gbk_ctx_create": (i32, [i32, pp]),
        "gbk_ctx_destroy": (i32, [VP]),
        "gbk_last_error": (C.c_char_p, [VP]),
        "gbk_ctx_stream": (VP, [VP]),
        "gbk_ctx_sync": (i32, [VP]),
        "gbk_ctx_set_param": (i32, [VP, C.c_char_p, i64]),
        "gbk_ctx_get_param": (i32, [VP, C.c_char_p, C.POINTER(i64)]),
        "gbk_ktime_read": (i32, [VP, C.POINTER(f64), C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]),
        "gbk_graph_upload": (i32, [VP, i64, i32, VP, VP, VP, VP, VP, VP, VP, pp]),
        "gbk_graph_build": (i32, [VP, i64, i64, VP, VP, VP, i32, i32, pp, C.POINTER(f64)]),
        "gbk_graph_info": (i32, [VP, C.POINTER(i64), C.POINTER(i64), C.POINTER(i32), C.POINTER(i32)]),
        "gbk_graph_download": (i32, [VP, VP, VP, VP, VP, VP, VP, VP, VP]),
        "gbk_graph_free": (i32, [VP, VP]),
        "gbk_relabel": (i32, [VP, VP, VP, pp, C.POINTER(RelabelStats)]),
        "gbk_mapping": (i32, [VP, VP, i32, VP, C.POINTER(f64)]),
        "gbk_relabel_by": (i32, [VP, VP, i32, pp, C.POINTER(RelabelStats), VP, C.POINTER(f64)]),
        "gbk_pr": (i32, [VP, VP, i32, i32, f64, i32, VP, C.POINTER(PrStats)]),
        "gbk_pr_noalloc": (i32, [VP, VP, i32, i32, f64, i32, VP, VP]),
        "gbk_pr_noalloc_max_bytes": (C.c_int64, []),
        "gbk_tier0_features": (i32, [VP, VP, C.c_int64, C.c_uint64, VP]),
        "gbk_bfs": (i32, [VP, VP, C.c_int32, i32, i32, VP, C.POINTER(BfsStats)]),
        "gbk_cc": (i32, [VP, VP, i32, VP, C.POINTER(CcStats)]),
        "gbk_sssp": (i32, [VP, VP, C.c_int32, C.c_int32, VP, C.POINTER(SsspStats)]),
        "gbk_bc": (i32, [VP, VP, VP, i32, VP, C.POINTER(BcStats)]),
        "gbk_bfs_batch": (i32, [VP, VP, VP, i32, VP, VP, VP, C.POINTER(BatchStats)]),
        "gbk_graph_features": (i32, [VP, VP, VP, i32]),
        "gbk_graph_union": (i32, [VP, VP, i32, VP, pp]),
        "gbk_cc_many": (i32, [VP, VP, VP, i32, i32, VP, C.POINTER(CcStats)]),
        "gbk_pr_many": (i32, [VP, VP, VP, i32, i32, i32, f64, i32, VP, VP, C.POINTER(PrStats)]),
        "gbk_mlp_create": (i32, [VP, i32, C.c_uint64, pp]),
        "gbk_mlp_destroy": (i32, [VP, VP]),
        "gbk_mlp_num_params": (i64, [VP]),
        "gbk_mlp_get_params": (i32, [VP, VP, VP]),
        "gbk_mlp_set_params": (i32, [VP, VP, VP]),
        "gbk_mlp_forward": (i32, [VP, VP, VP, i64, VP]),
        "gbk_mlp_grad": (i32, [VP, VP, VP, VP, i64, VP, C.POINTER(f64)]),
        "gbk_mlp_train": (i32, [VP, VP, VP, VP, i64, i32, i32, C.c_float, C.c_uint64, VP]),
        "gbk_sssp_batch": (i32, [VP, VP, VP, i32, C.c_int32, VP, VP, VP, C.POINTER(BatchStats)]),
        "gbk_bc_batch": (i32, [VP, VP, VP, i32, VP, C.POINTER(BatchStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols() -> list:
    """Names declared in include/gbk.h (the C ABI surface)."""
    import re
    hdr = open(os.path.join(os.path.dirname(_PKG), "include", "gbk.h")).read()
    return sorted(set(re.findall(r"^(?:u?int64_t|int|const char \*|void \*)\s*\*?(gbk_\w+)\(", hdr, re.M)))


@dataclass
class DeviceGraph:
    ctx: "Context"
    handle: int
    n: int
    m: int
    directed: bool
    weighted: bool

    def download(self) -> dict:
        L = lib()
        out = {"off": np.empty(self.n + 1, np.int64), "nbr": np.empty(max(self.m, 1), np.int32),
               "org_ids": np.empty(max(self.n, 1), np.int32)}
        out["wgt"] = np.empty(max(self.m, 1), np.int32) if self.weighted else None
        if self.directed:
            out["in_off"] = np.empty(self.n + 1, np.int64)
        self.ctx._check(L.gbk_graph_download(self.ctx.h, self.handle, _ptr(out["off"]), _ptr(out["nbr"]),
                                             _ptr(out["wgt"]), _ptr(out.get("in_off")), None, None,
                                             _ptr(out["org_ids"])))
        if self.directed:
            mi = int(out["in_off"][-1])
            out["in_nbr"] = np.empty(max(mi, 1), np.int32)
            out["in_wgt"] = np.empty(max(mi, 1), np.int32) if self.weighted else None
            self.ctx._check(L.gbk_graph_download(self.ctx.h, self.handle, None, None, None, None,
                                                 _ptr(out["in_nbr"]), _ptr(out["in_wgt"]), None))
            out["in_nbr"] = out["in_nbr"][:mi]
            if out["in_wgt"] is not None:
                out["in_wgt"] = out["in_wgt"][:mi]
        out["nbr"] = out["nbr"][:self.m]
        if out["wgt"] is not None:
            out["wgt"] = out["wgt"][:self.m]
        out["org_ids"] = out["org_ids"][:self.n]
        return out

    def free(self):
        if self.handle:
            lib().gbk_graph_free(self.ctx.h, self.handle)
            self.handle = 0

    def __del__(self):
        try:
            if self.handle and self.ctx.h:
                self.free()
        except Exception:
            pass


def _a(x, dt):
    return None if x is None else np.ascontiguousarray(x, dt)


class Context:
    """One HIP stream on one device; every call is stream-ordered on it."""

    def __init__(self, device: int = 0):
        L = lib()
        h = VP()
        rc = L.gbk_ctx_create(device, C.byref(h))
        if rc != GBK_OK:
            raise GbkUnavailable(f"gbk_ctx_create({device}) failed with {rc}: no usable GPU")
        self.h = h

    def close(self):
        if self.h:
            lib().gbk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != GBK_OK:
            msg = lib().gbk_last_error(self.h)
            raise GbkError(rc, msg.decode() if msg else "")

    @property
    def stream(self) -> int:
        return lib().gbk_ctx_stream(self.h)

    def sync(self):
        self._check(lib().gbk_ctx_sync(self.h))

    def set_param(self, name: str, value: int):
        self._check(lib().gbk_ctx_set_param(self.h, name.encode(), int(value)))

    def get_param(self, name: str) -> int:
        v = C.c_int64(0)
        self._check(lib().gbk_ctx_get_param(self.h, name.encode(), C.byref(v)))
        return v.value

    def ktime(self, kernel: int):
        """Selects the kernel (KT_*) whose launches are timed one by one."""
        self.set_param("ktime", kernel)

    def ktime_read(self) -> dict:
        """Reads and resets the per-launch timer."""
        ms = C.c_double(0)
        n, rows, edges = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        self._check(lib().gbk_ktime_read(self.h, C.byref(ms), C.byref(n), C.byref(rows), C.byref(edges)))
        return {"ms": ms.value, "launches": n.value, "rows": rows.value, "edges": edges.value}

    # ------------------------------------------------------------- graphs
    def upload(self, n, off, nbr, wgt=None, in_off=None, in_nbr=None, in_wgt=None,
               org_ids=None, directed=False) -> DeviceGraph:
        off, nbr, wgt = _a(off, np.int64), _a(nbr, np.int32), _a(wgt, np.int32)
        in_off, in_nbr, in_wgt = _a(in_off, np.int64), _a(in_nbr, np.int32), _a(in_wgt, np.int32)
        org_ids = _a(org_ids, np.int32)
        h = VP()
        self._check(lib().gbk_graph_upload(self.h, int(n), int(bool(directed)), _ptr(off), _ptr(nbr),
                                           _ptr(wgt), _ptr(in_off), _ptr(in_nbr), _ptr(in_wgt),
                                           _ptr(org_ids), C.byref(h)))
        return self._wrap(h)

    def build(self, u, v, w=None, n: int = -1, symmetrize=True, keep_self=False):
        """Device CSR construction from an edge list (returns graph, device ms)."""
        u, v, w = _a(u, np.int32), _a(v, np.int32), _a(w, np.int32)
        h = VP()
        ms = C.c_double(0)
        self._check(lib().gbk_graph_build(self.h, int(n), len(u), _ptr(u), _ptr(v), _ptr(w),
                                          int(bool(symmetrize)), int(bool(keep_self)), C.byref(h),
                                          C.byref(ms)))
        return self._wrap(h), ms.value

    def upload_graph(self, g) -> DeviceGraph:
        """Uploads any object with the oracle/host Graph attributes."""
        return self.upload(g.n, g.off, g.nbr, g.wgt, g.in_off if g.directed else None,
                           g.in_nbr if g.directed else None, g.in_wgt if g.directed else None,
                           getattr(g, "org_ids", None), g.directed)

    def _wrap(self, h) -> DeviceGraph:
        n, m, d, w = C.c_int64(), C.c_int64(), C.c_int(), C.c_int()
        self._check(lib().gbk_graph_info(h, C.byref(n), C.byref(m), C.byref(d), C.byref(w)))
        return DeviceGraph(self, h.value, n.value, m.value, bool(d.value), bool(w.value))

    # ------------------------------------------------------------ kernels
    def relabel(self, g: DeviceGraph, new_ids):
        ids = _a(new_ids, np.int32)
        if len(ids) != g.n:
            raise ValueError("mapping length differs from the vertex count")
        h = VP()
        st = RelabelStats()
        self._check(lib().gbk_relabel(self.h, g.handle, _ptr(ids), C.byref(h), C.byref(st)))
        return self._wrap(h), st

    def mapping(self, g: DeviceGraph, algo: int):
        """Ordering -o algo (0, 2, 5, 6 or 7) made on the device: (new_ids, ms)."""
        ids = np.zeros(g.n, np.int32)
        ms = C.c_double(0)
        self._check(lib().gbk_mapping(self.h, g.handle, int(algo), _ptr(ids), C.byref(ms)))
        return ids, ms.value

    def relabel_by(self, g: DeviceGraph, algo: int):
        """Ordering -o algo made and applied on the device: (graph, stats,
        new_ids, mapping ms)."""
        ids = np.zeros(g.n, np.int32)
        h = VP()
        st = RelabelStats()
        ms = C.c_double(0)
        self._check(lib().gbk_relabel_by(self.h, g.handle, int(algo), C.byref(h), C.byref(st), _ptr(ids),
                                         C.byref(ms)))
        return self._wrap(h), st, ids, ms.value

    def pr(self, g: DeviceGraph, variant=PR_GS, max_iters=20, tol=1e-4, fixed=False, download=True):
        out = np.empty(g.n, np.float32) if download else None
        st = PrStats()
        self._check(lib().gbk_pr(self.h, g.handle, variant, max_iters, tol, int(fixed), _ptr(out),
                                 C.byref(st)))
        return out, st

    def bfs(self, g: DeviceGraph, source: int, alpha=0, beta=0, download=True):
        out = np.empty(g.n, np.int32) if download else None
        st = BfsStats()
        self._check(lib().gbk_bfs(self.h, g.handle, int(source), alpha, beta, _ptr(out), C.byref(st)))
        return out, st

    def cc(self, g: DeviceGraph, variant=CC_AFFOREST, download=True):
        out = np.empty(g.n, np.int32) if download else None
        st = CcStats()
        self._check(lib().gbk_cc(self.h, g.handle, variant, _ptr(out), C.byref(st)))
        return out, st

    def sssp(self, g: DeviceGraph, source: int, delta: int = 1, download=True):
        out = np.empty(g.n, np.int32) if download else None
        st = SsspStats()
        self._check(lib().gbk_sssp(self.h, g.handle, int(source), int(delta), _ptr(out), C.byref(st)))
        return out, st

    def bc(self, g: DeviceGraph, sources: Sequence[int], download=True):
        s = np.ascontiguousarray(sources, np.int32)
        out = np.empty(g.n, np.float32) if download else None
        st = BcStats()
        self._check(lib().gbk_bc(self.h, g.handle, _ptr(s), len(s), _ptr(out), C.byref(st)))
        return out, st

    def bfs_batch(self, g: DeviceGraph, sources: Sequence[int], download=True):
        """Depths (sources x n, -1 unreached), per-source reached counts and
        out-edge sums, stats."""
        s = np.ascontiguousarray(sources, np.int32)
        depth = np.empty((len(s), g.n), np.int32) if download else None
        reached = np.empty(len(s), np.int64)
        edges = np.empty(len(s), np.int64)
        st = BatchStats()
        self._check(lib().gbk_bfs_batch(self.h, g.handle, _ptr(s), len(s), _ptr(depth), _ptr(reached),
                                        _ptr(edges), C.byref(st)))
        return depth, reached, edges, st

    def sssp_batch(self, g: DeviceGraph, sources: Sequence[int], delta: int = 1, download=True):
        """Distances (sources x n, INT32_MAX/2 unreached), per-source reached
        counts and out-edge sums, stats."""
        s = np.ascontiguousarray(sources, np.int32)
        dist = np.empty((len(s), g.n), np.int32) if download else None
        reached = np.empty(len(s), np.int64)
        edges = np.empty(len(s), np.int64)
        st = BatchStats()
        self._check(lib().gbk_sssp_batch(self.h, g.handle, _ptr(s), len(s), int(delta), _ptr(dist),
                                         _ptr(reached), _ptr(edges), C.byref(st)))
        return dist, reached, edges, st

    def bc_batch(self, g: DeviceGraph, sources: Sequence[int], download=True):
        """Brandes scores over all sources (normalised by the maximum, as
        bc()), 64 sources per traversal; stats.rounds counts BFS levels and
        stats.edges_reached the per-source reached out-edges (TEPS = 2x)."""
        s = np.ascontiguousarray(sources, np.int32)
        out = np.empty(g.n, np.float32) if download else None
        st = BatchStats()
        self._check(lib().gbk_bc_batch(self.h, g.handle, _ptr(s), len(s), _ptr(out), C.byref(st)))
        return out, st

    # ------------------------------------------------ many small graphs
    def graph_union(self, graphs):
        """One device graph holding the given graphs side by side, and the
        vertex base of each part (count + 1 entries)."""
        arr = (VP * len(graphs))(*[g.handle for g in graphs])
        vbase = np.empty(len(graphs) + 1, np.int64)
        h = VP()
        self._check(lib().gbk_graph_union(self.h, arr, len(graphs), _ptr(vbase), C.byref(h)))
        return self._wrap(h), vbase

    def cc_many(self, g: DeviceGraph, vbase, variant=CC_AFFOREST, download=True):
        vbase = np.ascontiguousarray(vbase, np.int64)
        out = np.empty(g.n, np.int32) if download else None
        st = CcStats()
        self._check(lib().gbk_cc_many(self.h, g.handle, _ptr(vbase), len(vbase) - 1, variant, _ptr(out),
                                      C.byref(st)))
        return out, st

    def pr_many(self, g: DeviceGraph, vbase, variant=PR_GS, max_iters=20, tol=1e-4, fixed=False, download=True):
        vbase = np.ascontiguousarray(vbase, np.int64)
        out = np.empty(g.n, np.float32) if download else None
        iters = np.empty(len(vbase) - 1, np.int32)
        st = PrStats()
        self._check(lib().gbk_pr_many(self.h, g.handle, _ptr(vbase), len(vbase) - 1, variant, max_iters, tol,
                                      int(fixed), _ptr(out), _ptr(iters), C.byref(st)))
        return out, iters, st

    # -------------------------------------------------- selector support
    def features(self, g: DeviceGraph) -> np.ndarray:
        """GBK_NUM_FEATURES graph statistics computed on the device."""
        out = np.empty(NUM_FEATURES, np.float32)
        rc = lib().gbk_graph_features(self.h, g.handle, _ptr(out), NUM_FEATURES)
        if rc < 0:
            self._check(rc)
        return out

    def tier0_features(self, g: DeviceGraph, sample_size: int = 0, seed: int = 0) -> dict:
        """The adaptive selector's Tier-0 features, computed on the device
        (the names and order of graphbrew_amd.host.tier0_features)."""
        from .host import TIER0_NAMES
        out = np.zeros(7, np.float64)
        self._check(lib().gbk_tier0_features(self.h, g.handle, int(sample_size), int(seed), _ptr(out)))
        return dict(zip(TIER0_NAMES, out.tolist()))

    def mlp(self, in_dim: int, seed: int = 0) -> "Mlp":
        return Mlp(self, in_dim, seed)


NUM_FEATURES = 12
FEATURE_NAMES = ("log2_n", "log2_m", "mean_degree", "log2_max_degree", "degree_cv", "hub_edge_share",
                 "clustering", "log_span", "near64", "near4096", "ws_l2_log2", "ws_mall_log2")


class Mlp:
    """fp32 MLP in -> 64 -> 64 -> 1 (ReLU) living on the device; forward,
    gradient and Adam training run in the library's own HIP kernels."""

    def __init__(self, ctx: Context, in_dim: int, seed: int = 0):
        self.ctx = ctx
        self.in_dim = in_dim
        h = VP()
        ctx._check(lib().gbk_mlp_create(ctx.h, int(in_dim), int(seed), C.byref(h)))
        self.h = h.value
        self.num_params = int(lib().gbk_mlp_num_params(self.h))

    def close(self):
        if self.h:
            lib().gbk_mlp_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _x(self, X):
        X = np.ascontiguousarray(X, np.float32)
        if X.ndim != 2 or X.shape[1] != self.in_dim:
            raise ValueError(f"expected rows x {self.in_dim} features")
        return X

    @property
    def params(self) -> np.ndarray:
        p = np.empty(self.num_params, np.float32)
        self.ctx._check(lib().gbk_mlp_get_params(self.ctx.h, self.h, _ptr(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.num_params:
            raise ValueError("parameter count differs")
        self.ctx._check(lib().gbk_mlp_set_params(self.ctx.h, self.h, _ptr(p)))

    def forward(self, X) -> np.ndarray:
        X = self._x(X)
        out = np.empty(len(X), np.float32)
        self.ctx._check(lib().gbk_mlp_forward(self.ctx.h, self.h, _ptr(X), len(X), _ptr(out)))
        return out

    def grad(self, X, y):
        X = self._x(X)
        y = np.ascontiguousarray(y, np.float32)
        g = np.empty(self.num_params, np.float32)
        loss = C.c_double(0)
        self.ctx._check(lib().gbk_mlp_grad(self.ctx.h, self.h, _ptr(X), _ptr(y), len(X), _ptr(g),
                                           C.byref(loss)))
        return loss.value, g

    def train(self, X, y, epochs=100, batch=256, lr=1e-3, seed=0) -> np.ndarray:
        X = self._x(X)
        y = np.ascontiguousarray(y, np.float32)
        curve = np.empty(epochs, np.float64)
        self.ctx._check(lib().gbk_mlp_train(self.ctx.h, self.h, _ptr(X), _ptr(y), len(X), int(epochs),
                                            int(batch), float(lr), int(seed), _ptr(curve)))
        return curve
