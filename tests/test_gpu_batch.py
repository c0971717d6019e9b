"""Multi-source batches against one-source-at-a-time oracles (GPU).

Bars: BFS depths and SSSP distances of every source bit-exact to the
oracle's single-source BFS / Dijkstra (every delta); per-source reached
counts and out-edge sums equal to those recomputed from the oracle's
answer.  Batches cover 1, odd, exactly 64 and more than 64 sources (two
chunks), repeated sources, unreachable parts and directed graphs.
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


def _star(leaves, weighted=False):
    u = np.zeros(leaves, np.int32)
    v = np.arange(1, leaves + 1, dtype=np.int32)
    w = (np.arange(leaves, dtype=np.int32) % 251 + 1) if weighted else None
    return O.make_graph(u, v, w)


def _directed(n=3000, m=30000, seed=5, weighted=False):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m).astype(np.int32)
    v = (rng.zipf(1.4, m) % n).astype(np.int32)
    w = rng.integers(1, 256, m).astype(np.int32) if weighted else None
    return O.make_graph(u, v, w, symmetrize=False, n=n)


def _sources(g, k, seed=0):
    rng = np.random.default_rng(seed)
    return rng.integers(0, g.n, k).astype(np.int32)


def _check_counts(g, vals, unreached, reached, edges):
    deg = np.diff(g.off)
    for j, row in enumerate(vals):
        ok = row != unreached
        assert reached[j] == int(ok.sum())
        assert edges[j] == int(deg[ok].sum())


BFS_GRAPHS = {
    "kron13": lambda: O.kron(13),
    "uniform12": lambda: O.kron(12, uniform=True),
    "star5000": lambda: _star(5000),
    "path": lambda: _fixture("path.el"),
    "disconnected": lambda: _fixture("disconnected.el"),
    "directed": lambda: _directed(),
}


@pytest.mark.parametrize("name", list(BFS_GRAPHS))
@pytest.mark.parametrize("k", [1, 7, 64, 70])
def test_bfs_batch_depths(ctx, name, k):
    g = BFS_GRAPHS[name]()
    dg = ctx.upload_graph(g)
    srcs = _sources(g, k, seed=k)
    srcs[-1] = srcs[0]  # a repeated source
    depth, reached, edges, st = ctx.bfs_batch(dg, srcs)
    for j, s in enumerate(srcs):
        ref = O.bfs_depths(g, int(s))
        assert np.array_equal(depth[j], ref), (name, k, j, int(s))
    _check_counts(g, depth, -1, reached, edges)
    assert st.edges_reached == int(edges.sum())
    assert st.rounds >= 1
    dg.free()


def test_bfs_batch_matches_single_source_parents(ctx):
    # the single-source kernel's parents define the same depths
    g = O.kron(12)
    dg = ctx.upload_graph(g)
    srcs = O.source_picker(g, 16)[1]
    depth, _, edges, _ = ctx.bfs_batch(dg, srcs)
    for j, s in enumerate(srcs):
        parent, st = ctx.bfs(dg, int(s))
        assert O.bfs_verify(g, int(s), parent)
        assert np.array_equal(parent >= 0, depth[j] >= 0)
        assert st.edges_reached == edges[j]
    dg.free()


SSSP_GRAPHS = {
    "kron12w": lambda: O.kron(12, weighted=True),
    "uniform11w": lambda: O.kron(11, uniform=True, weighted=True),
    "star3000w": lambda: _star(3000, weighted=True),
    "weighted_path": lambda: _fixture("weighted_path.wel", weighted=True, symmetrize=False),
    "directed_w": lambda: _directed(weighted=True),
}


@pytest.mark.parametrize("name", list(SSSP_GRAPHS))
@pytest.mark.parametrize("k", [1, 3, 17, 64, 65])
def test_sssp_batch_distances(ctx, name, k):
    g = SSSP_GRAPHS[name]()
    dg = ctx.upload_graph(g)
    srcs = _sources(g, k, seed=100 + k)
    srcs[-1] = srcs[0]
    refs = [O.dijkstra(g, int(s)) for s in srcs]
    for delta in (1, 32, 100000):
        dist, reached, edges, st = ctx.sssp_batch(dg, srcs, delta)
        for j in range(k):
            assert np.array_equal(dist[j], refs[j]), (name, k, delta, j)
        _check_counts(g, dist, O.K_DIST_INF, reached, edges)
        assert st.rounds >= 1 and st.buckets >= 1
    dg.free()


def test_sssp_batch_equals_single_source_counts(ctx):
    g = O.kron(13, weighted=True)
    dg = ctx.upload_graph(g)
    srcs = O.source_picker(g, 8)[1]
    _, reached, edges, _ = ctx.sssp_batch(dg, srcs, 64, download=False)
    for j, s in enumerate(srcs):
        _, st = ctx.sssp(dg, int(s), 64, download=False)
        assert st.reached == reached[j] and st.edges_reached == edges[j]
    dg.free()


def test_sssp_step_counts_reused_across_graphs_of_one_shape(ctx):
    """A call queues as many steps as the last one of its shape took before
    it first reads the state: a relabelled copy has the same shape but other
    step counts, and a path needs many more steps than the Kronecker graph
    before it; every call stays exact."""
    g = O.kron(12, weighted=True)
    ids = O.mapping_random(g.n)
    rg = O.relabel(g, ids)
    dg = ctx.upload_graph(g)
    drg, _ = ctx.relabel(dg, ids)
    srcs = _sources(g, 16, seed=7)
    for graph, dev, ss in ((g, dg, srcs), (rg, drg, ids[srcs]), (g, dg, srcs[::-1].copy())):
        refs = [O.dijkstra(graph, int(s)) for s in ss]
        dist, _, _, _ = ctx.sssp_batch(dev, ss, 1024)
        for j, s in enumerate(ss):
            assert np.array_equal(dist[j], refs[j])
        for s in ss[:3]:
            d1, _ = ctx.sssp(dev, int(s), 40)
            assert np.array_equal(d1, O.dijkstra(graph, int(s)))
    path = _fixture("weighted_path.wel", weighted=True, symmetrize=False)
    dp = ctx.upload_graph(path)
    for delta in (1, 1, 2):
        d1, _ = ctx.sssp(dp, 0, delta)
        assert np.array_equal(d1, O.dijkstra(path, 0))
    drg.free()
    dg.free()
    dp.free()


def test_sssp_batch_long_weighted_path(ctx):
    """64 sources on a weighted path: distances far above 65535, hundreds
    of buckets and rounds; every source equals Dijkstra."""
    n = 1500
    u = np.arange(n - 1, dtype=np.int32)
    g = O.make_graph(u, u + 1, np.full(n - 1, 200, np.int32))
    dg = ctx.upload_graph(g)
    srcs = np.arange(0, 64 * 23, 23, dtype=np.int32)
    dist, reached, edges, st = ctx.sssp_batch(dg, srcs, 1024)
    assert dist.max() > 65535
    for j, s in enumerate(srcs):
        assert np.array_equal(dist[j], O.dijkstra(g, int(s))), j
    _check_counts(g, dist, O.K_DIST_INF, reached, edges)
    dg.free()


BC_GRAPHS = {
    "kron13": lambda: O.kron(13),
    "uniform12": lambda: O.kron(12, uniform=True),
    "star3000": lambda: _star(3000),
    "path": lambda: _fixture("path.el"),
    "disconnected": lambda: _fixture("disconnected.el"),
    "directed": lambda: _directed(),
}


@pytest.mark.parametrize("name", list(BC_GRAPHS))
@pytest.mark.parametrize("k", [1, 4, 17, 64, 70])
def test_bc_batch_matches_brandes(ctx, name, k):
    """Scores over all sources against the serial Brandes oracle (1e-4
    relative and the reference's BCScoreWithinTolerance), the per-source
    reached out-edges against the oracle's BFS."""
    g = BC_GRAPHS[name]()
    dg = ctx.upload_graph(g)
    srcs = _sources(g, k, seed=200 + k)
    srcs[-1] = srcs[0]  # a repeated source counts twice, as in the reference's loop
    scores, st = ctx.bc_batch(dg, srcs)
    ref = O.bc_serial(g, srcs)
    bad = [i for i in range(g.n) if not O.bc_within_tolerance(float(ref[i]), float(scores[i]))]
    assert not bad, (name, k, bad[:5])
    np.testing.assert_allclose(scores, ref, rtol=1e-4, atol=1e-6)
    deg = np.diff(g.off)
    want = sum(int(deg[O.bfs_depths(g, int(s)) >= 0].sum()) for s in srcs)
    assert st.edges_reached == want
    assert st.rounds >= len(srcs) // 64
    dg.free()


@pytest.mark.parametrize("name", ["kron13", "star3000", "disconnected", "directed"])
@pytest.mark.parametrize("k", [4, 64])
@pytest.mark.parametrize("pull", [0, 2])
def test_bc_batch_directions_match_brandes(ctx, name, k, pull):
    """Forward levels forced top-down (0) or bottom-up from in-neighbours at
    every level (2) give the serial Brandes scores and the same reached
    out-edges; the default switches by frontier edges per level."""
    g = BC_GRAPHS[name]()
    dg = ctx.upload_graph(g)
    srcs = _sources(g, k, seed=300 + k)
    ctx.set_param("bc_pull", pull)
    try:
        scores, st = ctx.bc_batch(dg, srcs)
    finally:
        ctx.set_param("bc_pull", 1)
    ref = O.bc_serial(g, srcs)
    np.testing.assert_allclose(scores, ref, rtol=1e-4, atol=1e-6)
    deg = np.diff(g.off)
    assert st.edges_reached == sum(int(deg[O.bfs_depths(g, int(s)) >= 0].sum()) for s in srcs)
    dg.free()


def test_bc_batch_equals_one_source_runs(ctx):
    g = O.kron(14)
    dg = ctx.upload_graph(g)
    srcs = O.source_picker(g, 8)[1]
    batch, _ = ctx.bc_batch(dg, srcs)
    single, _ = ctx.bc(dg, srcs)
    np.testing.assert_allclose(batch, single, rtol=1e-4, atol=1e-6)
    dg.free()


def test_batch_rejects_bad_input(ctx):
    g = O.kron(8, weighted=True)
    dg = ctx.upload_graph(g)
    with pytest.raises(gbk.GbkError):
        ctx.bfs_batch(dg, [0, g.n])
    with pytest.raises(gbk.GbkError):
        ctx.sssp_batch(dg, [-1], 4)
    with pytest.raises(gbk.GbkError):
        ctx.sssp_batch(dg, [0], 0)
    with pytest.raises(gbk.GbkError):
        ctx.bc_batch(dg, [g.n])
    gu = ctx.upload_graph(O.kron(8))
    with pytest.raises(gbk.GbkError):
        ctx.sssp_batch(gu, [0], 4)
    gu.free()
    dg.free()


@pytest.mark.parametrize("k", [3, 8, 20, 64])
def test_batch_counts_without_download_equal(ctx, k):
    """Without the per-source arrays the reached counts and reached edges come
    from a streaming count over the vertex-major array; they equal the
    transposing path's and the oracle's."""
    g = O.kron(12, weighted=True)
    dg = ctx.upload_graph(g)
    src = O.source_picker(g, k)[1].astype(np.int32)
    d1, r1, e1, s1 = ctx.bfs_batch(dg, src)
    d2, r2, e2, s2 = ctx.bfs_batch(dg, src, download=False)
    assert d2 is None and np.array_equal(r1, r2) and np.array_equal(e1, e2)
    assert s1.edges_reached == s2.edges_reached == int(e1.sum())
    deg = np.diff(g.off)
    for j in range(k):
        reached = d1[j] >= 0
        assert r1[j] == reached.sum() and e1[j] == deg[reached].sum()
    _, r3, e3, s3 = ctx.sssp_batch(dg, src, 64)
    _, r4, e4, s4 = ctx.sssp_batch(dg, src, 64, download=False)
    assert np.array_equal(r3, r4) and np.array_equal(e3, e4) and s3.edges_reached == s4.edges_reached
    _, b1 = ctx.bc_batch(dg, src)
    _, b2 = ctx.bc_batch(dg, src, download=False)
    assert b1.edges_reached == b2.edges_reached


@pytest.mark.parametrize("group", [1, 8, 32, 64])
def test_bfs_batch_bottom_up_groups(ctx, group):
    """Every lanes-per-vertex width of the batch's bottom-up step gives the
    oracle's depths for every source (the step only changes who ORs which
    in-neighbour words)."""
    g = O.kron(13)
    dg = ctx.upload_graph(g)
    srcs = _sources(g, 64, seed=400)
    ctx.set_param("msbfs_group", group)
    try:
        depth = ctx.bfs_batch(dg, srcs)[0]
    finally:
        ctx.set_param("msbfs_group", 32)
    for j, s in enumerate(srcs):
        assert np.array_equal(depth[j], O.bfs_depths(g, int(s))), (group, j)
    dg.free()


def test_bc_batch_first_call_on_new_context():
    """A BC batch as the first call on a fresh context: the counter staging it
    reads after the per-source count pass stays valid (the staging area is
    never reallocated), and the scores equal serial Brandes."""
    g = O.kron(12)
    c = gbk.Context(0)
    try:
        dg = c.upload_graph(g)
        srcs = _sources(g, 4, seed=17)
        scores, st = c.bc_batch(dg, srcs)
        np.testing.assert_allclose(scores, O.bc_serial(g, srcs), rtol=1e-4, atol=1e-6)
        assert st.rounds > 0
        dg.free()
    finally:
        c.close()
