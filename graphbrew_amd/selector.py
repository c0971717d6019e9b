"""Learned ordering selector: predicts each candidate ordering's end-to-end
GPU time (mapping + relabel + reuse x kernel) from device-computed graph
statistics, the kernel and the reuse count, and picks the fastest.

The reference's deployed selector is a frozen threshold rule
(reorder/reorder_adaptive.h:302-333, ShouldUseAllKernelLowReuseRule): when
it matches it asks for the Leiden-Gorder composition, otherwise for Rabbit
order, or DBG (-o 5) in builds without Rabbit (:370-383).  Neither Leiden
nor Rabbit is produced here (those reorderers stay CPU code whose mapping
files enter through -o 13), so among the eight cheap orderings this build
produces the rule's deployable pick is DBG; it is the fallback of this
selector when no trained model is available.

Candidates: -o 0..7 (original, random, sort, hubsort, hubcluster, dbg,
hubsortdbg, hubclusterdbg).  Timings come from this library's own kernels on
seeded synthetic graph families (RMAT with varied a/b/c, stochastic block
models, 2D lattices with random rewiring, power-law configuration models),
measured by `collect()`; the network (graphbrew_amd.gbk.Mlp) is trained on
the GPU by the library's kernels.
"""
from __future__ import annotations

import json
import math
import os
import time
from dataclasses import dataclass, field
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import gbk
from . import host as H

ORDERINGS = tuple(range(8))
ORDER_NAMES = ("original", "random", "sort", "hubsort", "hubcluster", "dbg", "hubsortdbg", "hubclusterdbg")
KERNELS = ("pr", "pr_spmv", "bfs", "cc", "cc_sv", "sssp", "bc")
REUSE = (1, 2)
FALLBACK = 5  # DBG: the frozen rule's deployable pick here
# orderings the drivers make on the device (gbk_mapping / gbk_relabel_by):
# their mapping cost is device time; the others are made on the host
DEVICE_MAPPED = (1, 2, 5, 6, 7)
IN_DIM = gbk.NUM_FEATURES + len(ORDERINGS) + len(KERNELS) + 1
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_model.json")
MODEL_TEXT_PATH = os.path.splitext(MODEL_PATH)[0] + ".txt"


# ----------------------------------------------------------- graph families
def _finish(u, v, n, rng, permute):
    keep = u != v
    u, v = u[keep], v[keep]
    if permute:
        p = rng.permutation(n).astype(np.int32)
        u, v = p[u], p[v]
    return u.astype(np.int32), v.astype(np.int32)


def rmat(scale, ef, a, b, c, seed, permute=True):
    """R-MAT edge list (Chakrabarti et al.): each edge picks one quadrant per
    bit level with probabilities a, b, c, 1-a-b-c (host generator, seeded,
    parallel), self loops dropped."""
    n, m = 1 << scale, (1 << scale) * ef
    u, v = H.generate_rmat(scale, m, a, b, c, seed, permute)
    keep = u != v
    return (u[keep], v[keep]), n


def sbm(n, blocks, avg_degree, ratio, seed, permute=False):
    """Stochastic block model with contiguous blocks: each edge stays inside
    its source's block with probability ratio / (ratio + blocks - 1)."""
    rng = np.random.default_rng(seed)
    m = n * avg_degree // 2
    size = max(1, n // blocks)
    u = rng.integers(0, n, m)
    blk = np.minimum(u // size, blocks - 1)
    p_in = ratio / (ratio + blocks - 1)
    inside = rng.random(m) < p_in
    lo = blk * size
    hi = np.where(blk == blocks - 1, n, lo + size)
    v = np.where(inside, lo + (rng.random(m) * (hi - lo)).astype(np.int64), rng.integers(0, n, m))
    return _finish(u, v, n, rng, permute), n


def lattice(side, rewire, seed, permute=False):
    """side x side grid with 4-neighbour edges, a fraction rewired to random
    endpoints."""
    rng = np.random.default_rng(seed)
    n = side * side
    ids = np.arange(n).reshape(side, side)
    u = np.concatenate([ids[:, :-1].ravel(), ids[:-1, :].ravel()])
    v = np.concatenate([ids[:, 1:].ravel(), ids[1:, :].ravel()])
    sw = rng.random(len(v)) < rewire
    v = np.where(sw, rng.integers(0, n, len(v)), v)
    return _finish(u, v, n, rng, permute), n


def powerlaw(n, gamma, avg_degree, seed, permute=True):
    """Configuration model with a Zipf-like degree sequence of exponent gamma
    (stubs matched uniformly at random)."""
    rng = np.random.default_rng(seed)
    k = rng.pareto(gamma - 1.0, n) + 1.0
    k = np.maximum(1, np.round(k * avg_degree / k.mean())).astype(np.int64)
    k = np.minimum(k, n - 1)
    stubs = np.repeat(np.arange(n), k)
    rng.shuffle(stubs)
    if len(stubs) % 2:
        stubs = stubs[:-1]
    return _finish(stubs[0::2], stubs[1::2], n, rng, permute), n


def kron_specs() -> List[dict]:
    """The reference's own Kronecker generator (gapbs/generator.h, fixed
    seed) at scales 19-23 and degrees 8/16/32, without the bench graph
    (scale 22, degree 16): the family the bench graph comes from."""
    out = []
    for scale in (19, 20, 21, 22, 23):
        for ef in (8, 16, 32):
            if (scale, ef) == (22, 16) or (scale == 23 and ef != 16):
                continue
            out.append(dict(family="kron", scale=scale, ef=ef))
    return out


def corpus(count: int, seed0: int = 1000, min_scale: int = 12, max_scale: int = 16,
           skewed: bool = False, families: Sequence[str] = ("rmat", "sbm", "lattice", "powerlaw")) -> List[dict]:
    """Seeded graph specifications, cycling through the four families.
    skewed: R-MAT's b = c drawn near their upper bound (1 - a) / 2, as in the
    Kronecker generator's own 0.57 / 0.19 / 0.19, so heavy-tailed degrees are
    seen under permuted ids too and not only in generator order."""
    rng = np.random.default_rng(seed0)
    specs = []
    for i in range(count):
        seed = seed0 + i
        scale = int(rng.integers(min_scale, max_scale + 1))
        fam = families[i % len(families)]
        if fam == "rmat":
            a = float(rng.uniform(0.45, 0.65))
            lo = (1 - a) / 2 - 0.03 if skewed else 0.1
            b = c = float(rng.uniform(lo, (1 - a) / 2 - 0.005 if skewed else (1 - a) / 2 - 0.02))
            specs.append(dict(family=fam, scale=scale, ef=int(rng.choice([8, 16, 32])), a=a, b=b, c=c,
                              seed=seed, permute=bool(rng.random() < 0.7)))
        elif fam == "sbm":
            specs.append(dict(family=fam, n=1 << scale, blocks=int(2 ** rng.integers(4, 13)),
                              avg_degree=int(rng.choice([8, 16, 32])), ratio=float(rng.uniform(10, 100)),
                              seed=seed, permute=bool(rng.random() < 0.5)))
        elif fam == "lattice":
            specs.append(dict(family=fam, side=int(2 ** (scale / 2)), rewire=float(rng.uniform(0.0, 0.02)),
                              seed=seed, permute=bool(rng.random() < 0.5)))
        else:
            specs.append(dict(family=fam, n=1 << scale, gamma=float(rng.uniform(2.1, 3.0)),
                              avg_degree=int(rng.choice([8, 16])), seed=seed,
                              permute=bool(rng.random() < 0.7)))
    return specs


def make_edges(spec: dict):
    f = spec["family"]
    if f == "kron":
        u, v = H.generate_el(spec["scale"], spec["ef"], False)
        return (u, v), 1 << spec["scale"]
    if f == "rmat":
        return rmat(spec["scale"], spec["ef"], spec["a"], spec["b"], spec["c"], spec["seed"], spec["permute"])
    if f == "sbm":
        return sbm(spec["n"], spec["blocks"], spec["avg_degree"], spec["ratio"], spec["seed"], spec["permute"])
    if f == "lattice":
        return lattice(spec["side"], spec["rewire"], spec["seed"], spec["permute"])
    return powerlaw(spec["n"], spec["gamma"], spec["avg_degree"], spec["seed"], spec["permute"])


# ------------------------------------------------------------------ timing
def _median(xs):
    return float(np.median(np.asarray(xs, np.float64)))


def kernel_ms(ctx, name, g, gw, src, repeats=3) -> float:
    """Device time of one trial of a kernel as the bench binaries run it."""
    def once():
        if name == "pr":
            return ctx.pr(g, gbk.PR_GS, 20, 1e-4, False, download=False)[1].device_ms
        if name == "pr_spmv":
            return ctx.pr(g, gbk.PR_SPMV, 20, 1e-4, False, download=False)[1].device_ms
        if name == "bfs":
            return ctx.bfs(g, src, download=False)[1].device_ms
        if name == "cc":
            return ctx.cc(g, gbk.CC_AFFOREST, download=False)[1].device_ms
        if name == "cc_sv":
            return ctx.cc(g, gbk.CC_SV, download=False)[1].device_ms
        if name == "sssp":
            return ctx.sssp(gw, src, 64, download=False)[1].device_ms
        if name == "bc":
            return ctx.bc(g, [src], download=False)[1].device_ms
        raise ValueError(name)
    return _median([once() for _ in range(repeats)])


def measure_graph(ctx: gbk.Context, spec: dict, repeats: int = 3) -> dict:
    """Features of the graph in its given order, and per candidate ordering
    the mapping time (host), relabel time (device) and kernel times."""
    (u, v), n = make_edges(spec)
    w = ((np.arange(len(u), dtype=np.int64) * 2654435761) % 255 + 1).astype(np.int32)
    g, _ = ctx.build(u, v, None, n, True)
    gw, _ = ctx.build(u, v, w, n, True)
    feats = ctx.features(g)
    off = g.download()["off"]
    src_orig = int(H.source_picker(off, 1)[0][0])
    rows = {}
    for a in ORDERINGS:
        ids, map_ms = mapping_ms(ctx, g, off, a)
        if a == 0:
            # the graph is already in this order: nothing to map or relabel
            rg, rgw, rel_ms, rel_w_ms, map_ms = g, gw, 0.0, 0.0, 0.0
        else:
            rg, st = ctx.relabel(g, ids)
            rgw, st2 = ctx.relabel(gw, ids)
            rel_ms, rel_w_ms = st.device_ms, st2.device_ms
        src = int(ids[src_orig])
        times = {k: kernel_ms(ctx, k, rg, rgw, src, repeats) for k in KERNELS}
        rows[str(a)] = {"map_ms": map_ms, "relabel_ms": rel_ms, "relabel_w_ms": rel_w_ms, "kernel_ms": times}
        if a != 0:
            rg.free()
            rgw.free()
    out = {"spec": spec, "n": g.n, "m": g.m, "features": [float(x) for x in feats], "orderings": rows,
           "tier0": list(ctx.tier0_features(g).values())}
    g.free()
    gw.free()
    return out


def mapping_ms(ctx: gbk.Context, g, off, a: int, repeats: int = 3):
    """(new_ids, ms) of ordering a as the drivers make it: device time for
    DEVICE_MAPPED (median of repeats), host wall time otherwise."""
    if a in DEVICE_MAPPED:
        runs = [ctx.mapping(g, a) for _ in range(repeats)]
        return runs[0][0], _median([ms for _, ms in runs])
    t0 = time.perf_counter()
    ids = H.mapping(a, off)
    return ids, (time.perf_counter() - t0) * 1e3


def remap(ctx: gbk.Context, data: Sequence[dict], log=None) -> None:
    """Rebuilds every stored graph on the device and, in place, replaces the
    mapping times of DEVICE_MAPPED orderings by their device times (the host
    time is kept as map_host_ms) and stores the device Tier-0 features (the
    frozen rule's inputs)."""
    for i, row in enumerate(data):
        (u, v), n = make_edges(row["spec"])
        g, _ = ctx.build(u, v, None, n, True)
        if g.n != row["n"] or g.m != row["m"]:
            raise RuntimeError(f"graph {i} rebuilt with different size")
        row["tier0"] = list(ctx.tier0_features(g).values())
        for a in DEVICE_MAPPED:
            o = row["orderings"][str(a)]
            o.setdefault("map_host_ms", o["map_ms"])
            o["map_ms"] = mapping_ms(ctx, g, None, a)[1]
        g.free()
        if log:
            log(f"remap {i + 1}/{len(data)} n={row['n']} m={row['m']}")


def retime_relabel(ctx: gbk.Context, data: Sequence[dict], log=None, kernels: Sequence[str] = ()) -> None:
    """Rebuilds every stored graph on the device and, in place, replaces the
    relabel times of each ordering (unweighted and weighted CSR, device
    time) by those of the current relabel kernels, and the times of the
    listed kernels by the current ones; the other times are kept."""
    for i, row in enumerate(data):
        (u, v), n = make_edges(row["spec"])
        w = ((np.arange(len(u), dtype=np.int64) * 2654435761) % 255 + 1).astype(np.int32)
        g, _ = ctx.build(u, v, None, n, True)
        gw, _ = ctx.build(u, v, w, n, True)
        if g.n != row["n"] or g.m != row["m"]:
            raise RuntimeError(f"graph {i} rebuilt with different size")
        off = g.download()["off"]
        src_orig = int(H.source_picker(off, 1)[0][0])
        for a in ORDERINGS:
            o = row["orderings"][str(a)]
            if a == 0:
                for k in kernels:
                    o["kernel_ms"][k] = kernel_ms(ctx, k, g, gw, src_orig)
                continue
            ids = mapping_ms(ctx, g, off, a, repeats=1)[0]
            rel, rel_w = [], []
            for rep in range(3):
                rg, st = ctx.relabel(g, ids)
                rgw, st2 = ctx.relabel(gw, ids)
                rel.append(st.device_ms)
                rel_w.append(st2.device_ms)
                if rep < 2:  # the last relabelled pair is kept for the kernels
                    rg.free()
                    rgw.free()
            o["relabel_ms"], o["relabel_w_ms"] = _median(rel), _median(rel_w)
            for k in kernels:
                o["kernel_ms"][k] = kernel_ms(ctx, k, rg, rgw, int(ids[src_orig]))
            rg.free()
            rgw.free()
        g.free()
        gw.free()
        if log:
            log(f"relabel {i + 1}/{len(data)} n={row['n']} m={row['m']}")


def rule_decision(row: dict, kernel: str, reuse: int, llc: int = 0) -> bool:
    """The frozen all-kernel rule's decision on a stored graph (True: the
    Leiden-Gorder8 arm, which is not produced here; False: its fallback DBG)."""
    t0 = dict(zip(H.TIER0_NAMES, row["tier0"]))
    return H.rule_pick("allkernel-lowreuse-rule", kernel, reuse, row["n"], row["m"], t0, llc)[0]


def collect(ctx: gbk.Context, specs: Sequence[dict], path: Optional[str] = None, log=None) -> List[dict]:
    data = []
    for i, spec in enumerate(specs):
        t0 = time.time()
        data.append(measure_graph(ctx, spec))
        if log:
            log(f"graph {i + 1}/{len(specs)} {spec['family']} n={data[-1]['n']} m={data[-1]['m']} "
                f"{time.time() - t0:.1f}s")
        if path:
            with open(path, "w") as f:
                json.dump(data, f)
    return data


def e2e_ms(row: dict, kernel: str, reuse: int) -> float:
    """The reference's accounting (gapbs/builder.h:870-895): mapping +
    relabel + reuse x kernel; SSSP runs on the weighted CSR."""
    rel = row["relabel_w_ms"] if kernel == "sssp" else row["relabel_ms"]
    return row["map_ms"] + rel + reuse * row["kernel_ms"][kernel]


# ---------------------------------------------------------------- samples
def encode(features: Sequence[float], ordering: int, kernel: str, reuse: int, mean, std) -> np.ndarray:
    x = np.zeros(IN_DIM, np.float32)
    nf = gbk.NUM_FEATURES
    x[:nf] = (np.asarray(features, np.float64) - mean) / std
    x[nf + ordering] = 1.0
    x[nf + len(ORDERINGS) + KERNELS.index(kernel)] = 1.0
    x[-1] = reuse - 1.5
    return x


def samples(data: Sequence[dict], mean, std) -> Tuple[np.ndarray, np.ndarray]:
    """One row per (graph, kernel, reuse, candidate other than the original
    order); target = log(e2e(candidate) / e2e(original)), so the graph's own
    scale cancels and the network learns what an ordering changes."""
    X, y = [], []
    for g in data:
        for k in KERNELS:
            for r in REUSE:
                base = e2e_ms(g["orderings"]["0"], k, r)
                for a in ORDERINGS[1:]:
                    X.append(encode(g["features"], a, k, r, mean, std))
                    y.append(math.log(max(e2e_ms(g["orderings"][str(a)], k, r), 1e-6) / max(base, 1e-6)))
    return np.asarray(X, np.float32), np.asarray(y, np.float32)


MARGINS = (0.0, 0.01, 0.02, 0.05, 0.1, 0.2)


@dataclass
class Selector:
    params: Optional[np.ndarray] = None   # members x parameters (an ensemble)
    mean: Optional[np.ndarray] = None
    std: Optional[np.ndarray] = None
    margin: float = 0.0  # leave the original order only for a predicted gain above this
    meta: dict = field(default_factory=dict)

    def fit(self, ctx: gbk.Context, data: Sequence[dict], epochs=40, batch=256, lr=1e-3, seed=0, members=5):
        """Trains 'members' networks from different seeds on the GPU; their
        predictions are averaged (an ensemble varies less between graphs it
        has not seen than one network does)."""
        F = np.asarray([g["features"] for g in data], np.float64)
        self.mean = F.mean(0)
        self.std = F.std(0) + 1e-6
        X, y = samples(data, self.mean, self.std)
        ps, curves = [], []
        for k in range(members):
            net = ctx.mlp(IN_DIM, seed=seed + k)
            curves.append(net.train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed + k))
            ps.append(net.params)
            net.close()
        self.params = np.stack(ps)
        curve = np.mean(curves, axis=0)
        # the decision margin is chosen on the training graphs only
        self.margin = min(MARGINS, key=lambda mg: (evaluate(ctx, self, data, margin=mg)["model"]["mean_regret"], mg))
        self.meta = {"epochs": epochs, "batch": batch, "lr": lr, "rows": int(len(X)), "members": members,
                     "graphs": len(data), "final_loss": float(curve[-1]), "first_loss": float(curve[0]),
                     "margin": self.margin}
        return curve

    def predict(self, ctx: gbk.Context, features, kernel: str, reuse: int) -> np.ndarray:
        """Predicted end-to-end time of every candidate relative to the
        original order (1.0 for the original itself)."""
        X = np.stack([encode(features, a, kernel, reuse, self.mean, self.std) for a in ORDERINGS[1:]])
        net = ctx.mlp(IN_DIM)
        outs = []
        for p in np.atleast_2d(self.params):
            net.params = p
            outs.append(net.forward(X).astype(np.float64))
        net.close()
        return np.concatenate([[1.0], np.exp(np.mean(outs, axis=0))])

    def pick(self, ctx: gbk.Context, features, kernel: str, reuse: int, margin: Optional[float] = None) -> int:
        """The ordering to use; the frozen rule's deployable pick (DBG) when no
        trained model is available."""
        if self.params is None:
            return FALLBACK
        pred = self.predict(ctx, features, kernel, reuse)
        a = int(np.argmin(pred))
        mg = self.margin if margin is None else margin
        return a if pred[a] < 1.0 - mg else 0

    def save(self, path: str = MODEL_PATH):
        with open(path, "w") as f:
            json.dump({"in_dim": IN_DIM, "orderings": list(ORDER_NAMES), "kernels": list(KERNELS),
                       "params": [[float(f"{x:.7g}") for x in p] for p in np.atleast_2d(self.params)],
                       "mean": list(map(float, self.mean)),
                       "std": list(map(float, self.std)), "margin": self.margin, "meta": self.meta}, f)
        self.save_text(os.path.splitext(path)[0] + ".txt")

    def save_text(self, path: str = MODEL_TEXT_PATH):
        """The model for the C++ drivers (-o 14): a header line
        "gbk-selector 1 in_dim members margin features orderings kernels", the
        kernel names, the feature means, the feature deviations, then one line
        of parameters per ensemble member (gbk_mlp_set_params layout)."""
        P = np.atleast_2d(self.params)
        with open(path, "w") as f:
            f.write(f"gbk-selector 1 {IN_DIM} {len(P)} {self.margin:.9g} {gbk.NUM_FEATURES} "
                    f"{len(ORDERINGS)} {len(KERNELS)}\n")
            f.write(" ".join(KERNELS) + "\n")
            f.write(" ".join(f"{x:.17g}" for x in self.mean) + "\n")
            f.write(" ".join(f"{x:.17g}" for x in self.std) + "\n")
            for p in P:
                f.write(" ".join(f"{float(x):.9g}" for x in p) + "\n")

    @classmethod
    def load(cls, path: str = MODEL_PATH) -> "Selector":
        if not os.path.exists(path):
            return cls()
        with open(path) as f:
            d = json.load(f)
        if d.get("in_dim") != IN_DIM:
            return cls()
        return cls(np.atleast_2d(np.asarray(d["params"], np.float32)), np.asarray(d["mean"]),
                   np.asarray(d["std"]), float(d.get("margin", 0.0)), d.get("meta", {}))


def evaluate(ctx: gbk.Context, sel: "Selector", data: Sequence[dict], margin: Optional[float] = None) -> dict:
    """Regret (picked e2e / best measured e2e - 1) of the model, the frozen
    rule's deployable pick (DBG) and the original order, over all kernels,
    overall and per reuse count."""
    names = ("model", "fallback", "original")
    res = {r: {k: [] for k in names} for r in REUSE}
    picks = []
    leiden = 0
    llc = H.llc_bytes()
    for g in data:
        for k in KERNELS:
            for r in REUSE:
                e = np.array([e2e_ms(g["orderings"][str(a)], k, r) for a in ORDERINGS])
                best = e.min()
                p = sel.pick(ctx, g["features"], k, r, margin)
                picks.append(p)
                res[r]["model"].append(e[p] / best - 1)
                # the rule's deployable pick: DBG on its fallback branch, and DBG
                # again where it asks for Leiden-Gorder8, not produced here
                if "tier0" in g and rule_decision(g, k, r, llc):
                    leiden += 1
                res[r]["fallback"].append(e[FALLBACK] / best - 1)
                res[r]["original"].append(e[0] / best - 1)

    def summary(v):
        v = np.asarray(v)
        return {"mean_regret": float(v.mean()), "geomean_ratio": float(np.exp(np.mean(np.log1p(v)))),
                "within_10pct": float(np.mean(v <= 0.10))}

    out = {k: summary(res[1][k] + res[2][k]) for k in names}
    out["by_reuse"] = {str(r): {k: summary(res[r][k]) for k in names} for r in REUSE}
    out["picks"] = {ORDER_NAMES[a]: int(c) for a, c in zip(*np.unique(picks, return_counts=True))}
    out["cases"] = len(picks)
    out["rule_asks_leiden_gorder8"] = leiden
    return out


def split(data: Sequence[dict], held_out_every: int = 4):
    """Train / held-out split by seed: every 4th group of four consecutive
    graphs is held out (the families cycle with period four, so each family
    is on both sides)."""
    fam = 4
    held = [(i // fam) % held_out_every == held_out_every - 1 for i in range(len(data))]
    train = [g for g, h in zip(data, held) if not h]
    test = [g for g, h in zip(data, held) if h]
    return train, test


def main(argv=None):
    import argparse
    import sys
    p = argparse.ArgumentParser(description="collect timings / train / evaluate the ordering selector")
    p.add_argument("cmd", choices=["collect", "remeasure", "remap", "relabel", "train", "evaluate"])
    p.add_argument("--graphs", type=int, default=96)
    p.add_argument("--min-scale", type=int, default=12)
    p.add_argument("--max-scale", type=int, default=15)
    p.add_argument("--seed0", type=int, default=1000)
    p.add_argument("--skewed", action="store_true", help="collect: R-MAT b = c near (1 - a) / 2")
    p.add_argument("--families", default="rmat,sbm,lattice,powerlaw",
                   help="collect: families to cycle through, or 'kron' for the reference generator's graphs")
    p.add_argument("--corpus", default=os.path.join(os.path.dirname(MODEL_PATH), "selector_corpus.json"),
                   help="output of collect; for train / evaluate a comma-separated list to merge")
    p.add_argument("--model", default=MODEL_PATH)
    p.add_argument("--epochs", type=int, default=40)
    p.add_argument("--start", type=int, default=0, help="remeasure / remap / relabel: first graph of the corpus")
    p.add_argument("--count", type=int, default=-1, help="remeasure / remap / relabel: graphs (-1: to the end)")
    p.add_argument("--out", default="", help="remeasure / remap / relabel: output file")
    p.add_argument("--kernels", default="", help="relabel: kernels to re-time as well (comma list)")
    a = p.parse_args(argv)
    ctx = gbk.Context(0)
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    if a.cmd == "collect":
        specs = kron_specs() if a.families == "kron" else corpus(
            a.graphs, seed0=a.seed0, min_scale=a.min_scale, max_scale=a.max_scale, skewed=a.skewed,
            families=a.families.split(","))
        collect(ctx, specs, a.corpus, log)
        return 0
    if a.cmd == "remeasure":
        # the same graphs (their specs are stored with the timings) timed again
        # with the current kernels
        with open(a.corpus) as f:
            specs = [row["spec"] for row in json.load(f)]
        end = len(specs) if a.count < 0 else min(len(specs), a.start + a.count)
        collect(ctx, specs[a.start:end], a.out or a.corpus, log)
        return 0
    if a.cmd in ("remap", "relabel"):
        with open(a.corpus) as f:
            data = json.load(f)
        end = len(data) if a.count < 0 else min(len(data), a.start + a.count)
        if a.cmd == "remap":
            remap(ctx, data[a.start:end], log)
        else:
            retime_relabel(ctx, data[a.start:end], log, [k for k in a.kernels.split(",") if k])
        with open(a.out or a.corpus, "w") as f:
            json.dump(data, f)
        return 0
    data = []
    for path in a.corpus.split(","):
        with open(path) as f:
            data += json.load(f)
    train, test = split(data)
    if a.cmd == "train":
        sel = Selector()
        curve = sel.fit(ctx, train, epochs=a.epochs)
        sel.meta["held_out"] = evaluate(ctx, sel, test)
        sel.meta["train_eval"] = evaluate(ctx, sel, train)
        sel.save(a.model)
        log(f"loss {curve[0]:.4f} -> {curve[-1]:.4f}")
        print(json.dumps(sel.meta["held_out"]))
        return 0
    print(json.dumps(evaluate(ctx, Selector.load(a.model), test)))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
