#!/usr/bin/env python3
"""Benchmark of the GraphBrew hot path on one MI355X.

Workload (one "step"): on a seeded symmetrized Kronecker graph (scale 22,
edge factor 16, a=.57 b=.19 c=.19; SSSP weights U[1,255]) resident in HBM:
relabel under the chosen ordering, then PR pull (Gauss-Seidel, <=20 iters,
tol 1e-4), DO-BFS from 64 sources, CC Afforest, CC Shiloach-Vishkin,
delta-stepping SSSP from 64 sources and Brandes BC from 4 sources, sources
picked by the reference's SourcePicker (original ids, ordering-invariant).

value = whole-step traversed edges / step time, in GTEPS (each kernel's
TEPS numerator as the reference defines it).  Per-kernel GTEPS and the
fraction of the 8 TB/s HBM roofline (algorithmic bytes) are reported too;
'roofline' describes the kernel with the largest share of device time.
cpu_baseline times the CPU oracle (an OpenMP restatement of the reference
kernels) on a bounded sample of the same graph on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # bytes/s, MI355X HBM3E


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--scale", type=int, default=22)
    p.add_argument("--degree", type=int, default=16)
    p.add_argument("--uniform", action="store_true")
    p.add_argument("--order", default="identity", choices=["identity", "random", "sort"])
    p.add_argument("--bfs-sources", type=int, default=64)
    p.add_argument("--sssp-sources", type=int, default=64)
    p.add_argument("--bc-sources", type=int, default=4)
    p.add_argument("--delta", type=int, default=40, help="delta of one-source SSSP runs")
    p.add_argument("--batch-delta", type=int, default=1024,
                   help="bucket width of the 64-source SSSP batch (distances do not depend on it)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU sample work")
    p.add_argument("--only", default="", help="comma list of kernels to run (profiling)")
    p.add_argument("--verbose", action="store_true")
    p.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                   help="first copy MIB MiB device to device (calibrates HBM byte counters)")
    p.add_argument("--many-graphs", type=int, default=256,
                   help="small seeded R-MAT graphs (scales 10-16, ef 8-32) run as one batch; 0 skips")
    p.add_argument("--no-legs", action="store_true",
                   help="skip the other orderings and the one-source calls after the timed region")
    p.add_argument("--no-selector", action="store_true",
                   help="skip the learned-order vs identity end-to-end comparison")
    p.add_argument("--single-source", action="store_true",
                   help="run BFS and SSSP one source per call and BC one source per traversal, "
                        "instead of source batches")
    p.add_argument("--param", action="append", default=[], metavar="NAME=VALUE",
                   help="library tunable to set before the run (repeatable)")
    p.add_argument("--autotune", action="store_true",
                   help="pick delta / BU group / alpha-beta with the autotuner before timing")
    return p.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# Algorithmic HBM bytes (compulsory traffic).  o is the offset width of the
# definition: 4 B while the graph has fewer than 2^31 edges, else 8 B (the
# kernels read 8-byte offsets either way, so below 2^31 edges the figures do
# not credit the wider reads).  Work sets it from the bench graph.
OFF_W = 8


def offset_width(m):
    return 4 if m < 2 ** 31 else 8


def bytes_relabel(n, m, weighted=False):
    return 8 * m + (2 * OFF_W + 8) * n + (8 * m if weighted else 0)


def bytes_pr_iter(n, m):
    return 4 * m + (OFF_W + 16) * n


def bytes_bfs(n, m_r):
    return 4 * m_r + (OFF_W + 8) * n + n / 4


def bytes_cc(n, m):
    return 4 * m + (OFF_W + 16) * n


def bytes_sv_iter(n, m):
    return 4 * m + (OFF_W + 8) * n


def bytes_sssp(n, m_r):
    return 8 * m_r + (OFF_W + 8) * n


def bytes_bc(n, m_r):
    return 8 * m_r + m_r / 4 + (2 * OFF_W + 40) * n


class Workload:
    def __init__(self, args, ctx, gbk, H):
        self.args, self.ctx, self.gbk, self.H = args, ctx, gbk, H
        t0 = time.time()
        u, v = H.generate_el(args.scale, args.degree, args.uniform)
        log(f"generated {len(u)} edges in {time.time() - t0:.1f}s")
        self.g, ms = ctx.build(u, v, None, -1, True)
        global OFF_W
        OFF_W = offset_width(self.g.m)
        w = H.insert_weights(len(u))
        self.gw, ms2 = ctx.build(u, v, w, self.g.n, True)
        del u, v, w
        log(f"device CSR build: n={self.g.n} m={self.g.m} ({ms:.1f} ms, weighted {ms2:.1f} ms)")
        host_off = self.g.download()["off"] if args.order != "random" else None
        n = self.g.n
        if args.order == "identity":
            self.ids = H.mapping_identity(n)
        elif args.order == "random":
            self.ids = H.mapping_random(n)
        else:
            self.ids = H.mapping_sort(host_off)
        self.fingerprint = H.mapping_fingerprint(self.ids)
        self.kernels = [k for k in args.only.split(",") if k] or \
            ["relabel", "pr", "bfs", "cc", "cc_sv", "sssp", "bc"]
        self.rg = None
        self.rgw = None

    def step(self, rec):
        """One pass of the hot path; rec[name] accumulates (device ms, edges, bytes, calls)."""
        ctx, gbk, H, a = self.ctx, self.gbk, self.H, self.args

        def add(name, ms, edges, nbytes, calls=1):
            r = rec.setdefault(name, [0.0, 0, 0.0, 0])
            r[0] += ms
            r[1] += edges
            r[2] += nbytes
            r[3] += calls

        n = self.g.n
        if self.rg is not None:
            self.rg.free()
            self.rgw.free()
        self.rg, st = ctx.relabel(self.g, self.ids)
        if "relabel" in self.kernels:
            add("relabel", st.device_ms, self.g.m, bytes_relabel(n, self.g.m))
        self.rgw, stw = ctx.relabel(self.gw, self.ids)
        if "relabel" in self.kernels:
            add("relabel", stw.device_ms, self.gw.m, bytes_relabel(n, self.gw.m, True))
        g, gw = self.rg, self.rgw
        m = g.m
        if not hasattr(self, "srcs"):
            off = g.download()["off"]
            org = g.download()["org_ids"]
            self.srcs_orig, self.srcs = H.source_picker(off, max(a.bfs_sources, a.sssp_sources, a.bc_sources), org)
            self.deg = np.diff(off)
        if "pr" in self.kernels:
            _, st = ctx.pr(g, gbk.PR_GS, 20, 1e-4, False, download=False)
            add("pr", st.device_ms, m * st.iterations, bytes_pr_iter(n, m) * st.iterations, st.iterations)
        if "bfs" in self.kernels and a.single_source:
            for s in self.srcs[:a.bfs_sources]:
                _, st = ctx.bfs(g, int(s), download=False)
                add("bfs", st.device_ms, st.edges_reached, bytes_bfs(n, st.edges_reached))
        elif "bfs" in self.kernels:
            _, _, edges, st = ctx.bfs_batch(g, self.srcs[:a.bfs_sources], download=False)
            add("bfs", st.device_ms, st.edges_reached, sum(bytes_bfs(n, int(e)) for e in edges),
                len(edges))
        if "cc" in self.kernels:
            _, st = ctx.cc(g, gbk.CC_AFFOREST, download=False)
            add("cc", st.device_ms, m, bytes_cc(n, m))
        if "cc_sv" in self.kernels:
            _, st = ctx.cc(g, gbk.CC_SV, download=False)
            add("cc_sv", st.device_ms, m * st.iterations, bytes_sv_iter(n, m) * st.iterations, st.iterations)
        if "sssp" in self.kernels and a.single_source:
            for s in self.srcs[:a.sssp_sources]:
                _, st = ctx.sssp(gw, int(s), a.delta, download=False)
                add("sssp", st.device_ms, st.edges_reached, bytes_sssp(n, st.edges_reached))
        elif "sssp" in self.kernels:
            _, _, edges, st = ctx.sssp_batch(gw, self.srcs[:a.sssp_sources], a.batch_delta, download=False)
            add("sssp", st.device_ms, st.edges_reached, sum(bytes_sssp(n, int(e)) for e in edges),
                len(edges))
            scan = rec.setdefault("_sssp_scan", [0, 0, 0, 0])
            scan[0] += 1
            scan[1] += st.rounds
            scan[2] += st.rows_scanned
            scan[3] += st.edges_scanned
        if "bc" in self.kernels and a.single_source:
            _, st = ctx.bc(g, self.srcs[:a.bc_sources], download=False)
            add("bc", st.device_ms, 2 * st.edges_reached, bytes_bc(n, st.edges_reached), a.bc_sources)
        elif "bc" in self.kernels:
            k = len(self.srcs[:a.bc_sources])
            _, st = ctx.bc_batch(g, self.srcs[:a.bc_sources], download=False)
            # per-source figure summed over the sources (linear in m_r)
            add("bc", st.device_ms, 2 * st.edges_reached,
                bytes_bc(n, st.edges_reached) + (k - 1) * bytes_bc(n, 0), k)


# Kernels followed one launch at a time (HIP events recorded by the library
# on its stream around each launch) for the roofline of the dominant family.
def kt_table(gbk, single, msbfs_group=32):
    t = {
        "sssp": (gbk.KT_SSSP_RELAX, "sssp_relax+sssp_relax_heavy (one relaxation round)"),
        "pr": (gbk.KT_PR_ITER, "pr_seg_block + pr_block_close per row block (one Gauss-Seidel sweep; "
                               "pr_pull_block on graphs of at most 2^19 vertices)"),
        "bfs": (gbk.KT_BFS_STEP, "bfs_td_step+bfs_td_heavy | bfs_bu_step (one BFS step)"),
        "bc": (gbk.KT_BC_LEVEL, "bc_forward_level+bc_forward_heavy | bc_backward_level"),
    }
    if not single:
        t["sssp"] = (gbk.KT_SSSP_RELAX, "sb_relax<64>+sb_relax_heavy<64> "
                                        "(one relaxation round of a 64-source batch)")
        t["bfs"] = (gbk.KT_BFS_STEP, "msbfs_front_queue+msbfs_td+msbfs_td_heavy | "
                                     f"msbfs_bu<{msbfs_group}> (one level of a 64-source batch)")
        t["bc"] = (gbk.KT_BC_LEVEL, "bcb_forward+bcb_forward_heavy | bcb_backward+"
                                    "bcb_backward_heavy+bcb_backward_finish (one level of a source batch)")
    return t


def workload_name(args):
    return (f"{'uniform' if args.uniform else 'kronecker'}-scale{args.scale}-ef{args.degree}"
            f" relabel+pr+bfs{args.bfs_sources}+cc+cc_sv+sssp{args.sssp_sources}+bc{args.bc_sources}")


def launch_bytes(family, kt, work, rec):
    """Algorithmic bytes of one launch of the followed kernel.

    PR: one iteration reads 4 B per in-edge, the offsets (8 B per row), the
    contributions and scores (16 B per row): 4m + 24n.  SSSP: a relaxation
    round reads 8 B per scanned edge (neighbour + weight) and 16 B per
    non-stale frontier row (queue entry, offsets, distance); rows and edges
    are counted on the device outside the timed launches.  BFS and BC, and
    the 64-source batches: the per-source figures of all sources spread
    evenly over the launches (a batch reads each edge once for all its
    sources, so this counts the bytes separate runs would need)."""
    n, m = work.g.n, work.g.m
    if family == "pr":
        return float(bytes_pr_iter(n, m)), "exact"
    if family == "sssp" and work.args.single_source:
        return (8.0 * kt["edges"] + 16.0 * kt["rows"]) / max(kt["launches"], 1), "counted"
    return rec[family][2] / max(kt["launches"], 1), "per-source figure / launches"


def load_traffic(workload, family, single, field="bytes_per_launch"):
    """HBM bytes of the followed kernel, per launch or per call (field),
    measured with rocprofv3 PMC counters for this workload and mode
    (profiles/traffic.json)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    e = t.get(f"{workload}|{family}|{'single' if single else 'batched'}")
    return None if e is None else e.get(field)


def selector_e2e(work, gbk, H):
    """End-to-end time (mapping + relabel + reuse x kernel, the reference's
    accounting) of the learned pick, the original order and the frozen
    all-kernel rule on this workload's graph, for every kernel at reuse 1
    and 2.  Mapping times are charged as the drivers make them: on the
    device for the degree-keyed orderings, on the host otherwise.  The rule
    decides on the device Tier-0 features; where it asks for the
    Leiden-Gorder8 composition (a CPU reorderer not run here) its deployable
    stand-in is DBG, as on its fallback branch.  Runs after the timed region."""
    from graphbrew_amd import selector as S
    ctx = work.ctx
    sel = S.Selector.load()
    feats = ctx.features(work.g)
    t0 = ctx.tier0_features(work.g)
    off = work.g.download()["off"]
    src_orig = int(H.source_picker(off, 1)[0][0])
    picks = {k: {r: sel.pick(ctx, feats, k, r) for r in S.REUSE} for k in S.KERNELS}
    needed = sorted({0, S.FALLBACK} | {a for k in picks for a in picks[k].values()})
    rows = {}
    for a in needed:
        ids, map_ms = S.mapping_ms(ctx, work.g, off, a)
        if a == 0:
            g, gw, rel, relw, map_ms = work.g, work.gw, 0.0, 0.0, 0.0
        else:
            g, st = ctx.relabel(work.g, ids)
            gw, st2 = ctx.relabel(work.gw, ids)
            rel, relw = st.device_ms, st2.device_ms
        src = int(ids[src_orig])
        rows[a] = {"map_ms": map_ms, "relabel_ms": rel, "relabel_w_ms": relw,
                   "kernel_ms": {k: S.kernel_ms(ctx, k, g, gw, src, repeats=1) for k in S.KERNELS}}
        if a != 0:
            g.free()
            gw.free()
    out, ratios = {}, []
    for k in S.KERNELS:
        out[k] = {}
        for r in S.REUSE:
            learned = S.e2e_ms(rows[picks[k][r]], k, r)
            ident = S.e2e_ms(rows[0], k, r)
            hit, wsr = H.rule_pick("allkernel-lowreuse-rule", k, r, work.g.n, work.g.m, t0)
            e = {a: S.e2e_ms(rows[a], k, r) for a in rows}
            out[k][f"reuse{r}"] = {"pick": S.ORDER_NAMES[picks[k][r]], "learned_ms": round(learned, 3),
                                   "identity_ms": round(ident, 3),
                                   "frozen_rule": ("leiden_gorder8 (not produced here; dbg stands in)" if hit
                                                   else "dbg (fallback branch)"),
                                   "frozen_rule_ms": round(e[S.FALLBACK], 3),
                                   "best_of_these": S.ORDER_NAMES[min(e, key=e.get)],
                                   "learned_over_best_of_these": round(learned / min(e.values()), 4)}
            ratios.append(learned / ident)
    return {"model_graphs": sel.meta.get("graphs"), "margin": sel.margin,
            "tier0": {k: round(v, 6) for k, v in t0.items()},
            "learned_over_identity_geomean": round(float(np.exp(np.mean(np.log(ratios)))), 4),
            "kernels": out}


def many_graphs(ctx, gbk, count):
    """PR (Gauss-Seidel, <=20 iterations, tol 1e-4) and Afforest over
    'count' seeded R-MAT graphs of uneven size (scales 10-16, edge factor
    8-32, varied a/b/c), once as one union batch and once one graph per
    call; device times from the library's hipEvents."""
    from graphbrew_amd import workloads as WL
    graphs, edges = [], 0
    t0 = time.time()
    for u, v, n in WL.many_small_graphs(count):
        g, _ = ctx.build(u, v, None, n, True)
        graphs.append(g)
        edges += g.m
    gen_s = time.time() - t0
    ug, vbase = ctx.graph_union(graphs)
    out = {"graphs": count, "vertices": int(ug.n), "edges_directed": int(ug.m),
           "generation_s": round(gen_s, 1)}
    for name, run_many, run_one in (
            ("pr", lambda: ctx.pr_many(ug, vbase, gbk.PR_GS, 20, 1e-4, False, download=False)[2].device_ms,
             lambda g: ctx.pr(g, gbk.PR_GS, 20, 1e-4, False, download=False)[1].device_ms),
            ("cc", lambda: ctx.cc_many(ug, vbase, gbk.CC_AFFOREST, download=False)[1].device_ms,
             lambda g: ctx.cc(g, gbk.CC_AFFOREST, download=False)[1].device_ms)):
        run_many()
        batch_ms = float(np.median([run_many() for _ in range(3)]))
        t = time.time()
        one_ms = sum(run_one(g) for g in graphs)
        one_wall = (time.time() - t) * 1e3
        out[name] = {"batch_ms": round(batch_ms, 3), "one_by_one_device_ms": round(one_ms, 3),
                     "one_by_one_wall_ms": round(one_wall, 1),
                     "speedup_vs_one_by_one_wall": round(one_wall / batch_ms, 1)}
    for g in graphs:
        g.free()
    ug.free()
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(work, budget_s, kern, gpu_scores):
    """The CPU oracle (an OpenMP restatement of the reference kernels) on the
    same workload as the GPU step: both relabels from the same graphs, PR
    Gauss-Seidel to the same tolerance, Afforest, SV, and BFS / SSSP (the
    same delta) / BC from the same SourcePicker sources.  The per-source
    kernels run on a bounded sample of the sources and are scaled to the
    step's source counts.  value = the GPU step's traversed edges / the CPU
    step time.  Also the checker of the GPU run's PageRank: its L1 distance
    from the FP64 fixed point (the reference prints its own L1 change,
    pr.cc:175)."""
    import oracle as O
    a = work.args
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    gi, gwi = work.g.download(), work.gw.download()
    g0 = O.Graph(n=work.g.n, off=gi["off"], nbr=gi["nbr"])
    gw0 = O.Graph(n=work.gw.n, off=gwi["off"], nbr=gwi["nbr"], wgt=gwi["wgt"])
    del gi, gwi
    t = time.time()
    g = O.relabel(g0, work.ids)
    gw = O.relabel(gw0, work.ids)
    rel_s = time.time() - t
    del g0, gw0
    out = {"relabel": {"s": rel_s, "edges": 2 * g.m}}
    t = time.time()
    _, it, _ = O.pr_gs(g, max_iters=20, tol=1e-4, threads=threads)
    out["pr"] = {"s": time.time() - t, "edges": g.m * it, "iterations": it}
    t = time.time()
    O.cc_afforest(g, threads=threads)
    out["cc"] = {"s": time.time() - t, "edges": g.m}
    t = time.time()
    O.cc_sv(g, threads=threads)
    out["cc_sv"] = {"s": time.time() - t, "edges": None}
    deg = np.diff(g.off)
    t = time.time()
    O.bc(g, work.srcs[:a.bc_sources], threads=threads)
    out["bc"] = {"s": time.time() - t, "edges": None}
    spent = sum(v["s"] for v in out.values())
    # per-source kernels: sources one at a time until the budget is used
    bfs_s = sssp_s = 0.0
    k = 0
    delta = a.delta if a.single_source else a.batch_delta
    while k < max(a.bfs_sources, a.sssp_sources) and (k < 2 or spent + bfs_s + sssp_s < budget_s):
        s = int(work.srcs[k])
        t = time.time()
        O.bfs_do(g, s, threads=threads)
        bfs_s += time.time() - t
        t = time.time()
        O.sssp_delta(gw, s, delta, threads=threads)
        sssp_s += time.time() - t
        k += 1
    out["bfs"] = {"s": bfs_s * a.bfs_sources / k, "edges": None, "sampled_sources": k}
    out["sssp"] = {"s": sssp_s * a.sssp_sources / k, "edges": None, "sampled_sources": k}
    step_s = sum(v["s"] for v in out.values())
    step_edges = sum(kern[name]["edges_per_step"] for name in kern)
    per = {}
    for name, v in out.items():
        e = kern.get(name, {}).get("edges_per_step")
        per[name] = {"s_per_step": round(v["s"], 4),
                     "gteps": round(e / v["s"] / 1e9, 4) if e and v["s"] > 0 else None,
                     "gpu_speedup": round(v["s"] * 1e3 / kern[name]["ms_per_step"], 1) if name in kern else None}
        for extra in ("iterations", "sampled_sources"):
            if extra in v:
                per[name][extra] = v[extra]
    res = {"value": step_edges / step_s / 1e9, "unit": "GTEPS", "cores": threads, "kind": "port",
           "cpu": cpu_model(),
           "sample": (f"the scale-{a.scale} step on the same graphs: relabel x2, pr (gs, tol 1e-4), cc, cc_sv, "
                      f"bc x{a.bc_sources}; bfs and sssp (delta {delta}) on {k} of the sources, scaled to "
                      f"{a.bfs_sources} / {a.sssp_sources}; {step_s:.1f}s of CPU per step"),
           "kernels": per}
    if gpu_scores is not None:
        fp = O.pr_fp64(g)
        res["pr_l1_error_vs_fp64"] = float(np.abs(gpu_scores.astype(np.float64) - fp).sum())
    return res


def orderings_leg(work, gbk, orders=(1, 2, 5)):
    """The step's kernels once under other orderings made on the device
    (random, degree sort, DBG): mapping and relabel device time and each
    kernel's device ms and GTEPS, next to the identity step.  After the
    timed region."""
    from graphbrew_amd import selector as S
    ctx, a = work.ctx, work.args
    src_orig = work.srcs_orig
    out = {}
    for algo in orders:
        rg, st, ids, map_ms = ctx.relabel_by(work.g, algo)
        rgw, stw, _, _ = ctx.relabel_by(work.gw, algo)
        srcs = ids[src_orig]
        k = {}
        _, p = ctx.pr(rg, gbk.PR_GS, 20, 1e-4, False, download=False)
        k["pr"] = (p.device_ms, rg.m * p.iterations)
        _, _, _, b = ctx.bfs_batch(rg, srcs[:a.bfs_sources], download=False)
        k["bfs"] = (b.device_ms, b.edges_reached)
        _, c = ctx.cc(rg, gbk.CC_AFFOREST, download=False)
        k["cc"] = (c.device_ms, rg.m)
        _, c = ctx.cc(rg, gbk.CC_SV, download=False)
        k["cc_sv"] = (c.device_ms, rg.m * c.iterations)
        _, _, _, q = ctx.sssp_batch(rgw, srcs[:a.sssp_sources], a.batch_delta, download=False)
        k["sssp"] = (q.device_ms, q.edges_reached)
        _, bc = ctx.bc_batch(rg, srcs[:a.bc_sources], download=False)
        k["bc"] = (bc.device_ms, 2 * bc.edges_reached)
        step = st.device_ms + stw.device_ms + sum(ms for ms, _ in k.values())
        edges = rg.m + rgw.m + sum(e for _, e in k.values())
        out[S.ORDER_NAMES[algo]] = {
            "mapping_ms": round(map_ms, 3), "relabel_ms": round(st.device_ms + stw.device_ms, 3),
            "step_ms": round(step, 3), "gteps": round(edges / step / 1e6, 2),
            "kernels_ms": {n: round(ms, 3) for n, (ms, _) in k.items()}}
        rg.free()
        rgw.free()
    return out


def config1_leg(ctx, gbk, H, calls=5):
    """PageRank (Gauss-Seidel, tol 1e-4) on Kronecker scale 16, edge factor
    16, under the identity and under the reference's own Leiden-Gorder8
    composition and RCM-BNF mappings of that graph (files the reference's
    CPU pipeline wrote, tests/golden/mappings; -o 12 / -o 11 in the
    drivers).  Relabel and the median of `calls` PR calls per ordering, the
    mapping file's fingerprint with each.  After the timed region."""
    gold = os.path.join(ROOT, "tests", "golden", "mappings")
    u, v = H.generate_el(16, 16)
    g, _ = ctx.build(u, v, None, -1, True)
    out = {"workload": f"kronecker-scale16-ef16 pr-gs n={g.n} m={g.m}"}
    maps = {"identity": H.mapping_identity(g.n)}
    for tag in ("leiden_gorder8", "rcm_bnf"):
        path = os.path.join(gold, f"kron16_ef16_{tag}.so")
        if os.path.exists(path):
            maps[tag] = H.mapping_read(path)
    for tag, ids in maps.items():
        if len(ids) != g.n:
            out[tag] = {"error": "mapping does not fit the graph"}
            continue
        rg, rst = ctx.relabel(g, ids)
        runs = [ctx.pr(rg, gbk.PR_GS, 20, 1e-4, False, download=False)[1] for _ in range(calls)]
        ms = sorted(r.device_ms for r in runs)[len(runs) // 2]
        out[tag] = {"fingerprint": H.mapping_fingerprint(ids), "relabel_ms": round(rst.device_ms, 3),
                    "pr_ms": round(ms, 3), "pr_iterations": runs[0].iterations,
                    "pr_gteps": round(g.m * runs[0].iterations / ms / 1e6, 2)}
        rg.free()
    g.free()
    return out


def single_source_leg(work, gbk):
    """The kernels the drop-in drivers run one source at a time: BFS and
    SSSP from each of the step's sources (one call per source, delta of
    --delta) and BC with one traversal per source.  After the timed region."""
    ctx, a = work.ctx, work.args
    g, gw, n = work.rg, work.rgw, work.rg.n
    out = {}
    ms = edges = nbytes = 0.0
    for s in work.srcs[:a.bfs_sources]:
        _, st = ctx.bfs(g, int(s), download=False)
        ms += st.device_ms
        edges += st.edges_reached
        nbytes += bytes_bfs(n, st.edges_reached)
    out["bfs"] = {"ms": round(ms, 3), "sources": a.bfs_sources, "gteps": round(edges / ms / 1e6, 2),
                  "hbm_frac": round(nbytes / (ms / 1e3) / HBM_PEAK, 4)}
    ms = edges = nbytes = 0.0
    for s in work.srcs[:a.sssp_sources]:
        _, st = ctx.sssp(gw, int(s), a.delta, download=False)
        ms += st.device_ms
        edges += st.edges_reached
        nbytes += bytes_sssp(n, st.edges_reached)
    out["sssp"] = {"ms": round(ms, 3), "sources": a.sssp_sources, "delta": a.delta,
                   "gteps": round(edges / ms / 1e6, 2), "hbm_frac": round(nbytes / (ms / 1e3) / HBM_PEAK, 4)}
    _, st = ctx.bc(g, work.srcs[:a.bc_sources], download=False)
    out["bc"] = {"ms": round(st.device_ms, 3), "sources": a.bc_sources,
                 "gteps": round(2 * st.edges_reached / st.device_ms / 1e6, 2),
                 "hbm_frac": round(bytes_bc(n, st.edges_reached) / (st.device_ms / 1e3) / HBM_PEAK, 4)}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        return 2
    from graphbrew_amd import gbk
    from graphbrew_amd import host as H
    import torch

    ctx = gbk.Context(0)
    for kv in args.param:
        name, value = kv.split("=", 1)
        ctx.set_param(name, int(value))
    if args.copy_probe:
        nb = args.copy_probe << 20
        a = torch.empty(nb, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"copy probe: {nb} bytes read, {nb} bytes written")
        del a, b
    work = Workload(args, ctx, gbk, H)
    stream = torch.cuda.ExternalStream(ctx.stream)
    wrec = {}
    for w in range(args.warmup):
        work.step(wrec)
        log(f"warmup step {w} done")
    ctx.sync()
    torch.cuda.synchronize()
    tuned = None
    if args.autotune:
        from graphbrew_amd.autotune import Autotuner
        if work.rg is None:
            work.step({})
        tuner = Autotuner(ctx)
        t0 = time.time()
        srcs = work.srcs[:4]
        best_s = None
        if "sssp" in work.kernels and args.single_source:
            best_s = tuner.tune_sssp(work.rgw, srcs)
        elif "sssp" in work.kernels:
            best_s = tuner.tune_sssp_batch(work.rgw, work.srcs[:args.sssp_sources])
        best_b = tuner.tune_bfs(work.rg, srcs) if "bfs" in work.kernels else None
        if best_s and args.single_source:
            args.delta = best_s["delta"]
        elif best_s:
            args.batch_delta = best_s["delta"]
        if best_s:
            tuner.apply(best_s)
        if best_b:
            tuner.apply(best_b)
        tuned = {"sssp": best_s, "bfs": best_b, "seconds": round(time.time() - t0, 1),
                 "sources": len(srcs)}
        log(f"autotune: {tuned}")
        for row in tuner.log:
            log(f"  {row}")
    kts = kt_table(gbk, args.single_source, ctx.get_param("msbfs_group"))
    if wrec:
        family = max((f for f in wrec if f in kts), key=lambda f: wrec[f][0], default=None)
    else:
        family = next((f for f in ("sssp", "bc", "bfs", "pr") if f in work.kernels), None)
    if family is not None:
        ctx.ktime(kts[family][0])
        ctx.ktime_read()
    rec = {}
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.time()
    ev0.record(stream)
    for s in range(args.steps):
        work.step(rec)
        log(f"step {s} done")
    ev1.record(stream)
    ctx.sync()
    torch.cuda.synchronize()
    wall = time.time() - t0
    total_ms = ev0.elapsed_time(ev1)
    kt = ctx.ktime_read() if family is not None else None
    ctx.ktime(gbk.KT_NONE)
    n, m = work.g.n, work.g.m
    kern = {}
    total_edges = 0
    scan = rec.pop("_sssp_scan", None)
    for name, (ms, edges, nbytes, calls) in rec.items():
        secs = ms / 1e3
        total_edges += edges
        kern[name] = {"ms_per_step": ms / args.steps, "gteps": edges / secs / 1e9 if secs else 0.0,
                      "edges_per_step": edges / args.steps,
                      "gbps": nbytes / secs / 1e9 if secs else 0.0,
                      "hbm_frac": nbytes / secs / HBM_PEAK if secs else 0.0,
                      "calls_per_step": calls / args.steps}
        if name == "bfs" and not args.single_source:
            # the bit-parallel batch reads each edge once per level for all 64
            # sources: the per-source byte formula x 64 is a one-source-
            # equivalent rate, not traffic; the counted bytes of a batch call
            # (rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/traffic.json) give
            # the fraction of the HBM rate actually moved
            kern[name]["hbm_frac_one_source_equivalent"] = kern[name].pop("hbm_frac")
            per_call = load_traffic(workload_name(args), "bfs", False, "bytes_per_call")
            batch_calls = calls / args.bfs_sources  # one batch call carries every source
            kern[name]["hbm_frac"] = (per_call * batch_calls / secs / HBM_PEAK) if per_call and secs else None
            kern[name]["counted_bytes_per_batch_call"] = per_call
            kern[name]["hbm_frac_basis"] = "counted bytes per batch call" if per_call else "not counted"
    if scan and "sssp" in kern:
        # work the batch actually did: queued rows relaxed and their edges
        # walked (each edge once for all sources of the batch)
        kern["sssp"].update({"rounds_per_call": scan[1] / scan[0], "rows_scanned_per_call": scan[2] / scan[0],
                             "edges_scanned_per_call": scan[3] / scan[0],
                             "edges_scanned_per_us": scan[3] / (rec["sssp"][0] * 1e3)})
    value = total_edges / (total_ms / 1e3) / 1e9
    workload = workload_name(args)
    roof = None
    if family is not None and kt and kt["launches"]:
        per_launch, how = launch_bytes(family, kt, work, rec)
        avg_s = kt["ms"] / kt["launches"] / 1e3
        achieved = per_launch / avg_s / 1e9
        dom = max(rec, key=lambda f: rec[f][0])
        # counted traffic: per launch, or per call spread over this run's
        # launches (the number of blind no-op launches per call varies; they
        # read nothing)
        traffic = load_traffic(workload, family, args.single_source)
        per_call = load_traffic(workload, family, args.single_source, "bytes_per_call")
        if traffic is None and per_call is not None and family == "sssp":
            sources = args.sssp_sources
            calls = args.steps * (sources if args.single_source else -(-sources // 64))
            traffic = round(per_call * calls / kt["launches"])
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(achieved * 1e9 / HBM_PEAK, 4),
                "traffic": traffic,
                "kernel": kts[family][1], "family": family, "dominant_family": dom,
                "launches": kt["launches"], "avg_launch_us": round(kt["ms"] / kt["launches"] * 1e3, 2),
                "bytes_per_launch": round(per_launch), "bytes_basis": how,
                "launch_share_of_step": round(kt["ms"] / total_ms, 4)}
        if family == "sssp" and args.single_source:
            roof["rows_scanned"] = kt["rows"]
            roof["edges_scanned"] = kt["edges"]
    out = {
        "metric": "GTEPS (all five kernels + GPU relabel, aggregate) with %HBM roofline per kernel",
        "value": round(value, 4), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(total_ms / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded Kronecker generator of the reference, weights U[1,255])",
        "config": {"workload": workload,
                   "n": n, "m_directed": m, "order": args.order, "mapping_fingerprint": work.fingerprint,
                   "offset_width_bytes": OFF_W,
                   "delta": args.delta if args.single_source else args.batch_delta,
                   "batched": not args.single_source, "autotuned": tuned,
                   "params": {k: ctx.get_param(k) for k in ("heavy_degree", "bu_group", "bfs_alpha", "bfs_beta",
                                                            "relax_grid")}},
        "kernels": kern,
        "roofline": roof,
        "wall_s": round(wall, 2),
    }
    gpu_scores = None
    if "pr" in work.kernels:
        gpu_scores, pst = ctx.pr(work.rg, gbk.PR_GS, 20, 1e-4, False)
        out["pr_final_error"] = pst.final_error  # the L1 change of the last sweep (pr.cc:175)
    if not args.no_legs and len(work.kernels) == 7:
        out["orderings"] = orderings_leg(work, gbk)
        if not args.single_source:
            out["single_source"] = single_source_leg(work, gbk)
        out["config1"] = config1_leg(ctx, gbk, H)
    if not args.no_selector:
        out["selector_e2e"] = selector_e2e(work, gbk, H)
    if args.many_graphs > 0:
        out["many_graphs"] = many_graphs(ctx, gbk, args.many_graphs)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(work, args.cpu_budget, kern, gpu_scores)
        if "pr_l1_error_vs_fp64" in out["cpu_baseline"]:
            out["pr_l1_error_vs_fp64"] = out["cpu_baseline"].pop("pr_l1_error_vs_fp64")
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
