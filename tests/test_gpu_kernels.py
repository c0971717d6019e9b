"""GPU parity for BFS, CC, SSSP and BC against the CPU oracle (C ABI).

Bars: BFS -- reachability and depths bit-exact, parents pass the
reference's BFSVerifier; CC -- labels bit-exact (smallest id per component);
SSSP -- distances bit-exact; BC -- every score within the reference's
BCScoreWithinTolerance (1e-5 abs + 1e-5 rel) and 1e-4 relative of the
oracle's serial Brandes.
"""
import os

import numpy as np
import pytest

import oracle as O
from graphbrew_amd import gbk
from tests import bruteforce as B

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "graphs")


def _fixture(name, weighted=False, symmetrize=True):
    u, v, w = B.read_el(os.path.join(GOLD, name), weighted)
    return O.make_graph(u, v, w, symmetrize=symmetrize)


def _depths_from_parents(g, src, parent):
    d = O.bfs_depths(g, src)
    return d


def _star(leaves, weighted=False):
    u = np.zeros(leaves, np.int32)
    v = np.arange(1, leaves + 1, dtype=np.int32)
    w = (np.arange(leaves, dtype=np.int32) % 251 + 1) if weighted else None
    return O.make_graph(u, v, w)


def _directed(n=2000, m=20000, seed=9, weighted=False):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m).astype(np.int32)
    v = (rng.zipf(1.4, m) % n).astype(np.int32)
    w = rng.integers(1, 256, m).astype(np.int32) if weighted else None
    return O.make_graph(u, v, w, symmetrize=False, n=n)


BFS_GRAPHS = {
    "kron14": lambda: O.kron(14),
    "uniform13": lambda: O.kron(13, uniform=True),
    "star9000": lambda: _star(9000),
    "path": lambda: _fixture("path.el"),
    "disconnected": lambda: _fixture("disconnected.el"),
    "directed": lambda: _directed(),
}


@pytest.mark.parametrize("name", list(BFS_GRAPHS))
def test_bfs_parity(ctx, name):
    g = BFS_GRAPHS[name]()
    dg = ctx.upload_graph(g)
    _, srcs = O.source_picker(g, 6)
    for s in srcs:
        s = int(s)
        for alpha, beta in ((15, 18), (1, 1), (10 ** 9, 18)):
            parent, st = ctx.bfs(dg, s, alpha, beta)
            assert O.bfs_verify(g, s, parent), (name, s, alpha, beta)
            depth = O.bfs_depths(g, s)
            assert np.array_equal(parent >= 0, depth >= 0)
            # depth of each vertex via its parent chain equals BFS depth
            reached = np.flatnonzero(parent >= 0)
            assert np.all(depth[parent[reached]] == np.where(reached == s, 0, depth[reached] - 1))
            assert st.reached == len(reached)
            assert st.edges_reached == int(np.diff(g.off)[reached].sum())


def test_bfs_bottom_up_group_sizes(ctx):
    g = O.kron(13)
    dg = ctx.upload_graph(g)
    s = int(O.source_picker(g, 1)[1][0])
    saved = ctx.get_param("bu_group")
    try:
        for G in (1, 2, 4, 8, 16, 32, 64):
            ctx.set_param("bu_group", G)
            parent, st = ctx.bfs(dg, s)
            assert O.bfs_verify(g, s, parent)
            assert st.steps_bu > 0
    finally:
        ctx.set_param("bu_group", saved)


CC_GRAPHS = {
    "kron14": lambda: O.kron(14),
    "uniform12": lambda: O.kron(12, uniform=True),
    "disconnected": lambda: _fixture("disconnected.el"),
    "isolated": lambda: _fixture("isolated.el"),
    "star9000": lambda: _star(9000),
    "directed": lambda: _directed(),
    "directed_tiny": lambda: _fixture("tiny.el", symmetrize=False),
}


@pytest.mark.parametrize("name", list(CC_GRAPHS))
def test_cc_parity(ctx, name):
    g = CC_GRAPHS[name]()
    dg = ctx.upload_graph(g)
    ref = O.cc_afforest(g)
    a, st = ctx.cc(dg, gbk.CC_AFFOREST)
    assert np.array_equal(a, ref)
    assert O.cc_verify(g, a)
    sv, st2 = ctx.cc(dg, gbk.CC_SV)
    assert np.array_equal(sv, O.cc_sv(g))
    if not g.directed:
        assert np.array_equal(sv, ref)


SSSP_GRAPHS = {
    "kron13w": lambda: O.kron(13, weighted=True),
    "uniform12w": lambda: O.kron(12, uniform=True, weighted=True),
    "star9000w": lambda: _star(9000, weighted=True),
    "weighted_path": lambda: _fixture("weighted_path.wel", weighted=True, symmetrize=False),
    "weighted_tiny": lambda: _fixture("weighted.wel", weighted=True, symmetrize=False),
    "directed_w": lambda: _directed(weighted=True),
}


@pytest.mark.parametrize("name", list(SSSP_GRAPHS))
def test_sssp_parity(ctx, name):
    g = SSSP_GRAPHS[name]()
    dg = ctx.upload_graph(g)
    _, srcs = O.source_picker(g, 4)
    for s in srcs:
        s = int(s)
        ref = O.dijkstra(g, s)
        for delta in (1, 8, 32, 128, 100000):
            dist, st = ctx.sssp(dg, s, delta)
            assert np.array_equal(dist, ref), (name, s, delta)
        assert O.distance_fingerprint(g, dist) == O.distance_fingerprint(g, ref)
        assert st.reached == int((ref != O.K_DIST_INF).sum())


@pytest.mark.parametrize("name", ["kron13w", "uniform12w", "star9000w", "directed_w"])
@pytest.mark.parametrize("sssp_heavy", [256, 64])
def test_sssp_heavy_pass_push_and_pull(ctx, name, sssp_heavy):
    """The heavy pass pushed from the bucket's vertices (sssp_pull 0), pulled
    into the unsettled ones (2) and chosen by edge counts (1), each with and
    without the Bellman-Ford tail after a pull (sssp_tail), give the same
    exact distances; a row threshold of 64 sends the pull's big rows through
    the heavy-row kernel."""
    g = SSSP_GRAPHS[name]()
    dg = ctx.upload_graph(g)
    _, srcs = O.source_picker(g, 3)
    try:
        ctx.set_param("sssp_heavy", sssp_heavy)
        for s in srcs:
            s = int(s)
            ref = O.dijkstra(g, s)
            for mode in (0, 2, 1):
                ctx.set_param("sssp_pull", mode)
                for tail in (0, 1):
                    ctx.set_param("sssp_tail", tail)
                    for delta in (4, 32, 64):
                        dist, st = ctx.sssp(dg, s, delta)
                        assert np.array_equal(dist, ref), (name, s, mode, tail, delta)
                        assert st.reached == int((ref != O.K_DIST_INF).sum())
    finally:
        ctx.set_param("sssp_pull", 1)
        ctx.set_param("sssp_tail", 1)
        ctx.set_param("sssp_heavy", 256)


def test_sssp_weighted_path_closed_form(ctx):
    g = _fixture("weighted_path.wel", weighted=True, symmetrize=False)
    dist, _ = ctx.sssp(ctx.upload_graph(g), 0, 2)
    assert dist.tolist() == [0, 1, 3, 6, 10, 15]


def test_sssp_rejects_unweighted(ctx):
    dg = ctx.upload_graph(O.kron(8))
    with pytest.raises(gbk.GbkError):
        ctx.sssp(dg, 0, 1)


BC_GRAPHS = {
    "kron13": lambda: O.kron(13),
    "uniform12": lambda: O.kron(12, uniform=True),
    "star9000": lambda: _star(9000),
    "path": lambda: _fixture("path.el"),
    "tiny": lambda: _fixture("tiny.el"),
}


@pytest.mark.parametrize("name", list(BC_GRAPHS))
def test_bc_parity(ctx, name):
    g = BC_GRAPHS[name]()
    dg = ctx.upload_graph(g)
    _, srcs = O.source_picker(g, 4)
    scores, st = ctx.bc(dg, srcs)
    ref = O.bc_serial(g, srcs)
    bad = [i for i in range(g.n) if not O.bc_within_tolerance(float(ref[i]), float(scores[i]))]
    assert not bad, (name, bad[:5], scores[bad[:5]], ref[bad[:5]])
    np.testing.assert_allclose(scores, ref, rtol=1e-4, atol=1e-6)


def test_bc_parity_with_many_heavy_rows(ctx):
    """A heavy threshold of 64 sends most rows of every level through the
    chunked heavy path of both passes."""
    g = O.kron(13)
    dg = ctx.upload_graph(g)
    _, srcs = O.source_picker(g, 4)
    old = ctx.get_param("heavy_degree")
    ctx.set_param("heavy_degree", 64)
    try:
        scores, _ = ctx.bc(dg, srcs)
    finally:
        ctx.set_param("heavy_degree", old)
    ref = O.bc_serial(g, srcs)
    np.testing.assert_allclose(scores, ref, rtol=1e-4, atol=1e-6)


def test_bc_repeated_runs_all_match(ctx):
    """Regression: the forward's successor-bit aggregation read the next
    lane's word with a shuffle that only lanes 0..62 executed, so a run of
    bitmap word 0 ending at lane 62 was never written (a few percent of
    runs lost one successor edge).  Thirty runs with fresh uploads and
    other calls in between must all match."""
    g = O.kron(13)
    _, srcs = O.source_picker(g, 4)
    ref = O.bc_serial(g, srcs)
    tiny = _fixture("tiny.el")
    for rep in range(30):
        dt = ctx.upload_graph(tiny)
        ctx.bc(dt, [0, 1])
        dt.free()
        dg = ctx.upload_graph(g)
        scores, _ = ctx.bc(dg, srcs)
        np.testing.assert_allclose(scores, ref, rtol=1e-4, atol=1e-6, err_msg=f"run {rep}")
        dg.free()


def test_kernels_after_relabel_are_ordering_invariant(ctx):
    """Distances / components / reachability of original ids survive any
    relabelling (sssp.cc:92-127 fingerprints; bench/tests/test_relabel_weights.cc:167-206)."""
    g = O.kron(13, weighted=True)
    dg = ctx.upload_graph(g)
    ids = O.mapping_random(g.n)
    rg, _ = ctx.relabel(dg, ids)
    r = O.relabel(g, ids)
    o, i0 = O.source_picker(g, 2)
    o2, i1 = O.source_picker(r, 2)
    assert np.array_equal(o, o2)
    for a, b in zip(i0, i1):
        d0, _ = ctx.sssp(dg, int(a), 32)
        d1, _ = ctx.sssp(rg, int(b), 32)
        assert O.distance_fingerprint(g, d0) == O.distance_fingerprint(r, d1)
    c0, _ = ctx.cc(dg)
    c1, _ = ctx.cc(rg)
    assert len(np.unique(c0)) == len(np.unique(c1))
