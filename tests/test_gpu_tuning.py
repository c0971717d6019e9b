"""Per-launch kernel timer and autotuner (GPU).

The timer's counts are checked against quantities fixed by the algorithm:
PR with a fixed iteration count launches its pull kernel exactly that many
times; a delta-stepping run relaxes every reached vertex at least once, so
the counted frontier rows and edges are at least the reached vertices and
their out-edges (with equality for delta = infinity-like widths on a path
graph, where no vertex is ever improved twice).  The autotuner must return a
candidate from its grid whose output equals the default's.
"""
import numpy as np
import pytest

import oracle as O
from graphbrew_amd import gbk
from graphbrew_amd.autotune import Autotuner, DELTAS, BU_GROUPS, ALPHA_BETA, SSSP_HEAVY

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _timer_off(ctx):
    yield
    ctx.ktime(gbk.KT_NONE)
    ctx.ktime_read()


def test_ktime_pr_launch_count(ctx):
    g = O.kron(12)
    dg = ctx.upload_graph(g)
    ctx.ktime(gbk.KT_PR_ITER)
    ctx.ktime_read()
    _, st = ctx.pr(dg, gbk.PR_SPMV, 7, 1e-4, True, download=False)
    kt = ctx.ktime_read()
    assert st.iterations == 7
    assert kt["launches"] == 7
    assert 0 < kt["ms"] <= st.device_ms
    assert kt["rows"] == 0 and kt["edges"] == 0
    dg.free()


def test_ktime_off_records_nothing(ctx):
    g = O.kron(10)
    dg = ctx.upload_graph(g)
    ctx.ktime(gbk.KT_NONE)
    ctx.ktime_read()
    ctx.pr(dg, gbk.PR_GS, 3, 1e-4, True, download=False)
    kt = ctx.ktime_read()
    assert kt["launches"] == 0 and kt["ms"] == 0
    dg.free()


def test_ktime_sssp_counts_cover_reached(ctx):
    """Every reached vertex is a row of some light round; with the heavy
    pass pushed (sssp_pull 0) every out-edge of a reached vertex is walked
    too.  A pulled heavy pass walks the unsettled vertices' in-edges instead,
    so the edge bound holds for the push only."""
    g = O.kron(13, weighted=True)
    dg = ctx.upload_graph(g)
    src = int(np.argmax(np.diff(g.off)))
    try:
        for mode in (0, 1, 2):
            ctx.set_param("sssp_pull", mode)
            ctx.ktime(gbk.KT_SSSP_RELAX)
            ctx.ktime_read()
            dist, st = ctx.sssp(dg, src, 32)
            kt = ctx.ktime_read()
            assert np.array_equal(dist, O.sssp_delta(g, src, 32)), mode
            assert kt["launches"] >= st.relax_rounds >= 1
            assert kt["rows"] >= st.reached
            if mode == 0:
                assert kt["edges"] >= st.edges_reached
    finally:
        ctx.set_param("sssp_pull", 1)
        ctx.ktime(gbk.KT_NONE)
    dg.free()


def test_ktime_sssp_path_exact(ctx):
    # a weighted path 0-1-...-k: from vertex 0 every vertex is reached once
    # with its final distance, so rows == reached and edges == edges_reached
    k = 300
    u = np.arange(k, dtype=np.int32)
    v = u + 1
    w = np.full(k, 3, np.int32)
    g = O.make_graph(u, v, w)
    dg = ctx.upload_graph(g)
    ctx.ktime(gbk.KT_SSSP_RELAX)
    ctx.ktime_read()
    dist, st = ctx.sssp(dg, 0, 1 << 20)
    kt = ctx.ktime_read()
    assert np.array_equal(dist, 3 * np.arange(k + 1, dtype=np.int32))
    assert kt["rows"] == st.reached == k + 1
    assert kt["edges"] == st.edges_reached == 2 * k
    dg.free()


def test_ktime_bfs_and_bc_steps(ctx):
    g = O.kron(12)
    dg = ctx.upload_graph(g)
    ctx.ktime(gbk.KT_BFS_STEP)
    ctx.ktime_read()
    _, st = ctx.bfs(dg, 0, download=False)
    kt = ctx.ktime_read()
    assert kt["launches"] == st.steps_td + st.steps_bu
    ctx.ktime(gbk.KT_BC_LEVEL)
    ctx.ktime_read()
    _, bst = ctx.bc(dg, [0], download=False)
    kt = ctx.ktime_read()
    # forward levels: max_depth + 1 (the last one finds nothing); backward: max_depth + 1
    assert kt["launches"] == 2 * (bst.max_depth + 1)
    dg.free()


def test_ktime_rejects_unknown_kernel(ctx):
    with pytest.raises(gbk.GbkError):
        ctx.ktime(99)


def test_relax_grid_param_keeps_distances(ctx):
    g = O.kron(12, weighted=True)
    dg = ctx.upload_graph(g)
    ref = O.sssp_delta(g, 5, 16)
    saved = ctx.get_param("relax_grid")
    try:
        for grid in (64, 256, 2048):
            ctx.set_param("relax_grid", grid)
            assert np.array_equal(ctx.sssp(dg, 5, 16)[0], ref)
        with pytest.raises(gbk.GbkError):
            ctx.set_param("relax_grid", 4096)
    finally:
        ctx.set_param("relax_grid", saved)
    dg.free()


def test_autotuner_sssp_and_bfs(ctx, tmp_path):
    g = O.kron(12, weighted=True)
    dg = ctx.upload_graph(g)
    cache = str(tmp_path / "tune.json")
    t = Autotuner(ctx, repeats=1, cache=cache)
    srcs = O.source_picker(g, 3)[1]
    heavy = ctx.get_param("sssp_heavy")
    best = t.tune_sssp(dg, srcs)
    assert best["delta"] in DELTAS and best["ms"] > 0
    assert best["sssp_heavy"] in SSSP_HEAVY + (heavy,)
    assert ctx.get_param("sssp_heavy") == heavy  # the sweep leaves the context as it was
    assert min(r["ms"] for r in t.log if r["kernel"] == "sssp") == best["ms"]
    bb = t.tune_bfs(dg, srcs)
    assert bb["bu_group"] in BU_GROUPS and (bb["alpha"], bb["beta"]) in ALPHA_BETA
    # cached: a second tuner reads the file and does not sweep
    t2 = Autotuner(ctx, repeats=1, cache=cache)
    assert t2.tune_sssp(dg, srcs) == best and not t2.log
    saved = (ctx.get_param("bu_group"), ctx.get_param("bfs_alpha"), ctx.get_param("bfs_beta"))
    try:
        t.apply(bb)
        assert ctx.get_param("bu_group") == bb["bu_group"]
        for s in srcs:
            p, _ = ctx.bfs(dg, int(s))
            assert np.array_equal(p >= 0, O.bfs_depths(g, int(s)) >= 0)
    finally:
        ctx.set_param("bu_group", saved[0])
        ctx.set_param("bfs_alpha", saved[1])
        ctx.set_param("bfs_beta", saved[2])
    dg.free()


def test_autotuner_sssp_batch(ctx):
    from graphbrew_amd.autotune import BATCH_DELTAS
    g = O.kron(11, weighted=True)
    dg = ctx.upload_graph(g)
    t = Autotuner(ctx, repeats=1)
    srcs = O.source_picker(g, 8)[1]
    best = t.tune_sssp_batch(dg, srcs)
    assert best["delta"] in BATCH_DELTAS
    assert len([r for r in t.log if r["kernel"] == "sssp_batch"]) == len(BATCH_DELTAS)
    dist = ctx.sssp_batch(dg, srcs, best["delta"])[0]
    for j, s in enumerate(srcs):
        assert np.array_equal(dist[j], O.dijkstra(g, int(s)))
    dg.free()


def test_autotuner_pr_blocks(ctx):
    """Every block layout that converges is timed and its vector checked
    against the default's; the pick leaves the parameters restored."""
    from graphbrew_amd.autotune import PR_BLOCKS
    g = O.kron(13)
    dg = ctx.upload_graph(g)
    t = Autotuner(ctx, repeats=1)
    best = t.tune_pr(dg)
    assert best["pr_blocks"] in PR_BLOCKS and best["pr_inplace"] in (0, 1)
    assert 1 <= len([r for r in t.log if r["kernel"] == "pr"]) <= 2 * len(PR_BLOCKS)
    assert ctx.get_param("pr_blocks") == 0 and ctx.get_param("pr_inplace") == 0
    t.apply(best)
    try:
        s, st = ctx.pr(dg, gbk.PR_GS)
        assert st.final_error < 1e-4 and O.pr_residual(g, s) < 1e-4
    finally:
        ctx.set_param("pr_blocks", 0)
        ctx.set_param("pr_inplace", 0)
    dg.free()


def test_autotuner_bc_batch_direction(ctx):
    from graphbrew_amd.autotune import BC_PULL
    g = O.kron(12)
    dg = ctx.upload_graph(g)
    t = Autotuner(ctx, repeats=1)
    srcs = O.source_picker(g, 8)[1]
    best = t.tune_bc_batch(dg, srcs)
    assert (best["bc_pull"], best["bc_pull_ratio"]) in BC_PULL
    assert len([r for r in t.log if r["kernel"] == "bc_batch"]) == len(BC_PULL)
    assert ctx.get_param("bc_pull") == 1 and ctx.get_param("bc_pull_ratio") == 1
    t.apply(best)
    try:
        scores = ctx.bc_batch(dg, srcs)[0]
    finally:
        ctx.set_param("bc_pull", 1)
        ctx.set_param("bc_pull_ratio", 1)
    np.testing.assert_allclose(scores, O.bc_serial(g, srcs), rtol=1e-4, atol=1e-6)
    dg.free()


def test_autotuner_bfs_batch_group(ctx):
    """Every bottom-up width of the BFS batch gives the oracle's depths; the
    sweep restores the parameter and the pick applies."""
    from graphbrew_amd.autotune import MSBFS_GROUPS
    g = O.kron(13)
    dg = ctx.upload_graph(g)
    t = Autotuner(ctx, repeats=1)
    srcs = O.source_picker(g, 64)[1]
    best = t.tune_bfs_batch(dg, srcs)
    assert best["msbfs_group"] in MSBFS_GROUPS
    assert len([r for r in t.log if r["kernel"] == "bfs_batch"]) == len(MSBFS_GROUPS)
    assert ctx.get_param("msbfs_group") == 32
    t.apply(best)
    try:
        depth = ctx.bfs_batch(dg, srcs)[0]
    finally:
        ctx.set_param("msbfs_group", 32)
    for j in (0, 17, 63):
        assert np.array_equal(depth[j], O.bfs_depths(g, int(srcs[j])))
    dg.free()
