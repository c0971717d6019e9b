"""The C++ bench drivers keep the reference binaries' interface: getopt
flags, stdout lines (gapbs/util.h:60-68 formats) and exit codes.  CPU tests
cover argument handling; GPU tests run whole benchmarks with -v and compare
fingerprints with the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from graphbrew_amd import host as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "graphbrew_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden", "graphs")


def run(name, *args, timeout=120):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip("drivers not built")
    print("driver:", name, *args, flush=True)
    return subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=timeout)


def labels(out):
    d = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            d.setdefault(k.strip(), []).append(v.strip())
    return d


@pytest.mark.parametrize("name", ["pr", "pr_spmv", "bfs", "cc", "cc_sv", "sssp", "bc"])
def test_no_input_is_rejected(name):
    r = run(name)
    assert r.returncode == 255
    assert "No graph input specified" in r.stdout


def test_help_exits_zero():
    r = run("bfs", "-h")
    assert r.returncode == 0 and "breadth-first search" in r.stdout


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("order", ["0", "1", "2", "3", "4", "5", "6", "7"])
def test_bfs_driver_verifies(order):
    r = run("bfs", "-g", 12, "-n", 4, "-v", "-a", "-o", order)
    assert r.returncode == 0, r.stdout + r.stderr
    L = labels(r.stdout)
    assert L["Verification"] == ["PASS"] * 4
    assert len(L["Trial Time"]) == 4 and "Average Time" in L
    g = O.kron(12)
    orig, _ = O.source_picker(g, 4)
    assert [int(x) for x in L["Source Original"]] == orig.tolist()
    ids = {"0": np.arange(g.n, dtype=np.int32), "1": O.mapping_random(g.n),
           "2": O.mapping_sort(g), "3": O.mapping_hubsort(g.off), "4": O.mapping_hubcluster(g.off),
           "5": O.mapping_dbg(g), "6": O.mapping_hubsortdbg(g.off),
           "7": O.mapping_hubclusterdbg(g.off)}[order]
    assert L["Mapping Fingerprint"] == [O.mapping_fingerprint(ids)]
    assert re.search(r"^Average Time:\s+\d+\.\d{5}$", r.stdout, re.M)


@pytest.mark.gpu
def test_sssp_driver_fingerprints_are_ordering_invariant(tmp_path):
    fps = set()
    for order in ("0", "1", "2"):
        r = run("sssp", "-g", 11, "-n", 3, "-d", 16, "-v", "-o", order)
        assert r.returncode == 0, r.stdout + r.stderr
        L = labels(r.stdout)
        assert L["Verification"] == ["PASS"] * 3
        fps.add((L["Weight Checksum"][0], tuple(L["Distance Fingerprint"])))
    assert len(fps) == 1
    g = O.kron(11, weighted=True)
    assert fps.pop()[0] == O.weight_checksum(g)


@pytest.mark.gpu
def test_map_file_ordering(tmp_path):
    g = O.kron(11)
    ids = O.mapping_random(g.n)[::-1].copy()
    p = tmp_path / "m.lo"
    H.mapping_write(str(p), ids)
    r = run("cc_sv", "-g", 11, "-n", 1, "-v", "-o", f"13:{p}")
    assert r.returncode == 0, r.stdout + r.stderr
    L = labels(r.stdout)
    assert L["Mapping Fingerprint"] == [O.mapping_fingerprint(ids)]
    assert L["Verification"] == ["PASS"]


@pytest.mark.gpu
def test_pr_nonconvergence_matches_reference_and_exits_3():
    """Gauss-Seidel PR on the degree-sorted scale-12 graph does not reach the
    1e-4 L1 change within 20 iterations -- neither in the oracle's serial
    restatement of pr.cc nor on the GPU -- so PRVerifier fails and the
    harness exits 3 (gapbs/benchmark.h:57-64)."""
    g = O.kron(12)
    r = O.relabel(g, O.mapping_sort(g))
    s, iters, err = O.pr_gs(r)
    assert iters == 20 and err > 1e-4 and O.pr_residual(r, s) > 1e-4
    out = run("pr", "-g", 12, "-n", 1, "-v", "-o", 2)
    assert out.returncode == 3
    L = labels(out.stdout)
    assert L["Verification"] == ["FAIL"] and L["Verification Failure"] == ["pr"]
    assert float(L["Iterations"][0]) == 20.0


@pytest.mark.gpu
@pytest.mark.parametrize("name,extra", [("pr", ["-o", "1"]), ("pr_spmv", ["-o", "2"]), ("cc", ["-a"]),
                                        ("cc_sv", []), ("bc", ["-i", "2"])])
def test_kernel_drivers_verify(name, extra):
    r = run(name, "-g", 12, "-n", 2, "-v", *extra)
    assert r.returncode == 0, r.stdout + r.stderr
    assert labels(r.stdout)["Verification"] == ["PASS", "PASS"]


@pytest.mark.gpu
def test_file_input_and_components():
    r = run("cc", "-f", os.path.join(GOLD, "disconnected.el"), "-s", "-n", 1, "-v", "-a")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "There are 3 components" in r.stdout
    r = run("sssp", "-f", os.path.join(GOLD, "weighted_path.wel"), "-n", 1, "-r", 0, "-v")
    assert r.returncode == 0, r.stdout + r.stderr
    d = np.array([0, 1, 3, 6, 10, 15], np.int32)
    assert labels(r.stdout)["Distance Fingerprint"] == [H.distance_fingerprint(d)]


@pytest.mark.gpu
def test_unavailable_ordering_fails_cleanly():
    r = run("bfs", "-g", 10, "-n", 1, "-o", "12:leiden")
    assert r.returncode != 0 and "-o 13:<file>" in (r.stderr + r.stdout)


# ------------------------------------------------- serialized graph input
@pytest.mark.gpu
def test_drivers_read_serialized_graphs(tmp_path):
    """-f .sg / .wsg (ReadSerializedGraph, gapbs/reader.h:445-578): the CSR
    is taken as stored.  A .wsg of the generated scale-10 graph gives the
    same SSSP distance fingerprints as -g 10; BFS and CC verify on the .sg."""
    g = O.kron(10, weighted=True)
    sg, wsg = str(tmp_path / "k10.sg"), str(tmp_path / "k10.wsg")
    H.sg_write(sg, g.off, g.nbr)
    H.sg_write(wsg, g.off, g.nbr, wgt=g.wgt)
    for name in ("bfs", "cc", "pr"):
        r = run(name, "-f", sg, "-n", 2, "-v")
        assert r.returncode == 0, r.stdout + r.stderr
        assert labels(r.stdout)["Verification"] == ["PASS", "PASS"]
    a = run("sssp", "-f", wsg, "-n", 3, "-v", "-d", 2)
    b = run("sssp", "-g", 10, "-n", 3, "-v", "-d", 2)
    assert a.returncode == 0 and b.returncode == 0, a.stdout + b.stdout
    La, Lb = labels(a.stdout), labels(b.stdout)
    assert La["Verification"] == ["PASS"] * 3
    assert La["Distance Fingerprint"] == Lb["Distance Fingerprint"]


@pytest.mark.gpu
def test_driver_symmetrizes_directed_sg(tmp_path):
    """-s on a directed unweighted .sg with identity original ids rebuilds
    it symmetrized (gapbs/builder.h:741-790); CC then verifies."""
    rng = np.random.default_rng(8)
    u = rng.integers(0, 2000, 12000).astype(np.int32)
    v = rng.integers(0, 2000, 12000).astype(np.int32)
    d = O.make_graph(u, v, symmetrize=False, n=2000)
    p = str(tmp_path / "d.sg")
    H.sg_write(p, d.off, d.nbr, in_off=d.in_off, in_nbr=d.in_nbr)
    r = run("cc", "-f", p, "-s", "-n", 1, "-v")
    assert r.returncode == 0, r.stdout + r.stderr
    assert labels(r.stdout)["Verification"] == ["PASS"]


# ------------------------------------------- reference reorderers and -o 14
MAPS = os.path.join(ROOT, "tests", "golden", "mappings")
LEIDEN8 = "leiden:compose:sg_none:comm_size_desc:intra_gorder:gw8"


def run_env(name, *args, env=None, timeout=180):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip("drivers not built")
    e = dict(os.environ)
    e.pop("GBK_MAPPING_DIR", None)
    e.update(env or {})
    print("driver:", name, *args, flush=True)
    return subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=timeout, env=e)


@pytest.mark.gpu
@pytest.mark.parametrize("spec,tag", [("11:bnf", "rcm_bnf"), ("12:" + LEIDEN8, "leiden_gorder8")])
def test_reference_reorderer_mapping_files(spec, tag):
    """-o 11:bnf and -o 12:<Leiden-Gorder8> take the reference reorderers'
    mappings from GBK_MAPPING_DIR; the applied mapping is the file's
    (fingerprint) and BFS verifies on the relabelled graph."""
    r = run_env("bfs", "-g", 16, "-n", 2, "-v", "-o", spec, env={"GBK_MAPPING_DIR": MAPS})
    assert r.returncode == 0, r.stdout + r.stderr
    L = labels(r.stdout)
    ids = H.mapping_read(os.path.join(MAPS, f"kron16_ef16_{tag}.so"))
    assert L["Mapping Fingerprint"] == [H.mapping_fingerprint(ids)]
    assert L["Reorder Mapping File"] == [f"kron16_ef16_{tag}.so"]
    assert L["Verification"] == ["PASS", "PASS"]
    r = run_env("bfs", "-g", 16, "-n", 1, "-o", spec)
    assert r.returncode != 0 and "-o 13:<file>" in (r.stderr + r.stdout)


def _rule_expectation(kernel, reuse, rule="allkernel-lowreuse-rule"):
    g = O.kron(16)
    t0 = H.tier0_features(g.off, g.nbr)
    hit, wsr = H.rule_pick(rule, kernel, reuse, g.n, g.m, t0)
    return g, hit, wsr


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,reuse", [("bfs", 1), ("pr_spmv", 2), ("cc", 1), ("sssp", 3)])
def test_adaptive_frozen_rule(kernel, reuse):
    """-o 14:_:_:_:allkernel-lowreuse-rule:best-endtoend:<reuse> decides by
    the reference's predicate on the device Tier-0 features; its Leiden-Gorder8
    arm comes from the mapping file, else DBG stands in; the fallback is DBG."""
    g, hit, wsr = _rule_expectation(kernel, reuse)
    spec = f"14:_:_:_:allkernel-lowreuse-rule:best-endtoend:{reuse}"
    dbg = H.mapping_fingerprint(H.mapping_dbg(g.off))
    arm = H.mapping_fingerprint(H.mapping_read(os.path.join(MAPS, "kron16_ef16_leiden_gorder8_lowreuse.so")))
    for env, want in (({"GBK_MAPPING_DIR": MAPS}, arm if hit else dbg), ({}, dbg)):
        r = run_env(kernel, "-g", 16, "-n", 1, "-o", spec, env=env)
        assert r.returncode == 0, r.stdout + r.stderr
        L = labels(r.stdout)
        assert L["Selection"] == ["model=allkernel-lowreuse-rule criterion=best-endtoend"]
        assert float(L["Property WSR LLC"][0]) == pytest.approx(wsr, abs=6e-6)  # printed %3.5lf
        assert L["Adaptive Rule"] == ["allkernel-lowreuse-rule " + ("match" if hit else "fallback")]
        assert L["Mapping Fingerprint"] == [want], (env, hit)


@pytest.mark.gpu
def test_adaptive_budgeted_rule_and_learned_pick(ctx):
    """budgeted-rule follows ShouldUseBudgetedLeidenGorder; the default
    (perceptron) model is the learned selector, whose pick equals the Python
    selector's for the same graph, kernel and reuse."""
    g, hit, _ = _rule_expectation("pr", 1, "budgeted-rule")
    r = run_env("pr", "-g", 16, "-n", 1, "-o", "14:_:_:_:budgeted-rule:best-endtoend:1",
                env={"GBK_MAPPING_DIR": MAPS})
    assert r.returncode == 0, r.stdout + r.stderr
    L = labels(r.stdout)
    assert L["Adaptive Rule"] == ["budgeted-rule " + ("match" if hit else "fallback")]
    from graphbrew_amd import selector as S
    sel = S.Selector.load()
    dg = ctx.upload_graph(g)
    feats = ctx.features(dg)
    for kernel, reuse in (("bfs", 1), ("cc_sv", 2), ("pr", 2)):
        want = sel.pick(ctx, feats, kernel, reuse)
        r = run_env(kernel, "-g", 16, "-n", 1, "-o", f"14:_:_:_:perceptron:best-endtoend:{reuse}")
        assert r.returncode == 0, r.stdout + r.stderr
        L = labels(r.stdout)
        assert "Adaptive Learned Prediction" in L
        assert L["Adaptive Selected"] == [str(want)], (kernel, reuse)
    dg.free()


@pytest.mark.gpu
@pytest.mark.parametrize("spec,msg", [
    ("14:x", "AdaptiveOrder reserved fields must be empty"),
    ("14:_:_:_:allkernel-lowreuse-rule", "requires an explicit reuse count"),
    ("14:_:_:_:allkernel-lowreuse-rule:fastest-execution:1", "requires best-endtoend criterion"),
    ("14:_:_:_:dt:e2e:1", "offline-only"),
    ("14:_:_:_:perceptron:e2e:0", "reuse count must be positive"),
    ("14:_:_:_:perceptron:e2e:1:extra", "no longer accepts"),
    ("14:_:_:_:nosuch", "Unknown adaptive selection model"),
])
def test_adaptive_policy_errors(spec, msg):
    """ParseDeployableSelectionPolicy's refusals (reorder_adaptive.h:165-237)."""
    r = run_env("bfs", "-g", 10, "-n", 1, "-o", spec)
    assert r.returncode != 0 and msg in (r.stderr + r.stdout), r.stderr + r.stdout
