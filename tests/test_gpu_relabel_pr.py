"""GPU parity: relabel (bit-exact) and PageRank (tolerances stated below)
against the CPU oracle, through the C ABI."""
import numpy as np
import pytest

import oracle as O
from graphbrew_amd import gbk
from graphbrew_amd import host as H

pytestmark = pytest.mark.gpu


def _star(leaves: int, weighted=False):
    u = np.zeros(leaves, np.int32)
    v = np.arange(1, leaves + 1, dtype=np.int32)
    w = (np.arange(leaves, dtype=np.int32) % 251 + 1) if weighted else None
    return O.make_graph(u, v, w)


def _mappings(g):
    rng = np.random.default_rng(7)
    return {
        "identity": np.arange(g.n, dtype=np.int32),
        "random": O.mapping_random(g.n),
        "sort": O.mapping_sort(g),
        "reverse": np.arange(g.n - 1, -1, -1, dtype=np.int32),
        "shuffle": rng.permutation(g.n).astype(np.int32),
    }


def _assert_same_graph(dg, ref):
    got = dg.download()
    assert np.array_equal(got["off"], ref.off)
    assert np.array_equal(got["nbr"], ref.nbr)
    if ref.wgt is not None:
        assert np.array_equal(got["wgt"], ref.wgt)
    if ref.directed:
        assert np.array_equal(got["in_off"], ref.in_off)
        assert np.array_equal(got["in_nbr"], ref.in_nbr)
        if ref.in_wgt is not None:
            assert np.array_equal(got["in_wgt"], ref.in_wgt)
    assert np.array_equal(got["org_ids"], ref.org_ids)


def _class_rows(weighted=True):
    """Directed graph whose rows sit exactly at the sort classes' edges (64 |
    65, 128 | 129, 256 | 257, 512 | 513, 2048 | 2049 entries)."""
    rng = np.random.default_rng(11)
    n = 8192
    sizes = [64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 513, 1000, 2048, 2049, 6000]
    us, vs = [], []
    for i, k in enumerate(sizes * 4):
        us.append(np.full(k, i * 7, np.int32))
        vs.append(rng.choice(n, k, replace=False).astype(np.int32))
    u = np.concatenate(us)
    v = np.concatenate(vs)
    w = rng.integers(1, 256, len(u)).astype(np.int32) if weighted else None
    return O.make_graph(u, v, w, symmetrize=False, n=n)


def _crowded(weighted=False):
    """Big rows whose new ids crowd into a narrow part of [0, n): under the
    identity, row 0's 3001 neighbours are 1..3000 and 32000 (the row's first
    bucket split twice more), row 2^17's 6000 neighbours are random."""
    n = 1 << 18
    rng = np.random.default_rng(5)
    hub2 = rng.choice(np.arange(2, n), 6000, replace=False)
    u = np.concatenate([np.zeros(3001), np.full(6000, 1 << 17)]).astype(np.int32)
    v = np.concatenate([np.arange(1, 3001), [32000], hub2]).astype(np.int32)
    w = rng.integers(1, 256, len(u)).astype(np.int32) if weighted else None
    return O.make_graph(u, v, w, n=n)


GRAPHS = {
    "crowded": lambda: _crowded(),
    "crowded_w": lambda: _crowded(True),
    "classes_w": lambda: _class_rows(),
    "classes": lambda: _class_rows(False),
    "kron12": lambda: O.kron(12),
    "kron14w": lambda: O.kron(14, weighted=True),
    "uniform12": lambda: O.kron(12, uniform=True),
    "star5000": lambda: _star(5000),
    "star70000w": lambda: _star(70000, weighted=True),
}


@pytest.mark.parametrize("name", list(GRAPHS))
def test_relabel_bit_exact(ctx, name):
    g = GRAPHS[name]()
    dg = ctx.upload_graph(g)
    for mname, ids in _mappings(g).items():
        print(f"relabel {name} {mname}", flush=True)
        r, st = ctx.relabel(dg, ids)
        _assert_same_graph(r, O.relabel(g, ids))
        assert st.device_ms > 0
        r.free()


def test_relabel_directed_weighted(ctx):
    rng = np.random.default_rng(3)
    n, m = 3000, 40000
    u = rng.integers(0, n, m).astype(np.int32)
    v = (rng.zipf(1.5, m) % n).astype(np.int32)  # skewed in-degrees
    w = rng.integers(1, 256, m).astype(np.int32)
    g = O.make_graph(u, v, w, symmetrize=False, n=n)
    dg = ctx.upload_graph(g)
    for ids in _mappings(g).values():
        r, _ = ctx.relabel(dg, ids)
        _assert_same_graph(r, O.relabel(g, ids))


def _multi_edge_csr(n, edges):
    """CSR of the (u, v, w) edge list as given: repeated (u, v) pairs with
    different weights stay (an uploaded CSR need not be squished)."""
    u, v, w = (np.asarray(x, np.int32) for x in edges)
    order = np.lexsort((w, v, u))
    u, v, w = u[order], v[order], w[order]
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, u + 1, 1)
    return np.cumsum(off), v.copy(), w.copy()


def test_relabel_weighted_repeated_neighbours(ctx):
    """Rows that hold one neighbour several times (weights differ) sort by
    (neighbour, weight) in every row class: the 32-bit (id, index) sorts
    find the repeats and hand those rows to the 64-bit redo."""
    rng = np.random.default_rng(17)
    n = 5000
    us, vs, ws = [], [], []
    for row, k, reps in [(0, 3000, 3), (1, 500, 2), (2, 40, 4), (3, 9, 2), (4, 1500, 1), (5, 200, 1)]:
        base = rng.choice(n, k // reps, replace=False)
        vv = np.repeat(base, reps)[:k]
        us.append(np.full(len(vv), row))
        vs.append(vv)
        ws.append(rng.permutation(len(vv)) + 1)
    bu, bv = rng.integers(0, n, 20000), rng.integers(0, n, 20000)
    us.append(bu)
    vs.append(bv)
    ws.append(rng.integers(1, 256, 20000))
    edges = (np.concatenate(us), np.concatenate(vs), np.concatenate(ws))
    off, nbr, wgt = _multi_edge_csr(n, edges)
    in_off, in_nbr, in_wgt = _multi_edge_csr(n, (edges[1], edges[0], edges[2]))
    g = O.Graph(n=n, off=off, nbr=nbr, wgt=wgt, in_off=in_off, in_nbr=in_nbr, in_wgt=in_wgt, directed=True)
    dg = ctx.upload_graph(g)
    for mname, ids in _mappings(g).items():
        print(f"relabel repeated neighbours {mname}", flush=True)
        r, _ = ctx.relabel(dg, ids)
        _assert_same_graph(r, O.relabel(g, ids))
        r.free()


def test_relabel_chain_composes_org_ids(ctx):
    g = O.kron(11)
    a, b = O.mapping_random(g.n), O.mapping_sort(g)
    d1, _ = ctx.relabel(ctx.upload_graph(g), a)
    d2, _ = ctx.relabel(d1, b)
    ref = O.relabel(O.relabel(g, a), b)
    _assert_same_graph(d2, ref)


def test_relabel_rejects_bad_mapping(ctx):
    g = O.kron(8)
    dg = ctx.upload_graph(g)
    bad = np.arange(g.n, dtype=np.int32)
    bad[3] = bad[4]
    with pytest.raises(gbk.GbkError) as e:
        ctx.relabel(dg, bad)
    assert e.value.code == gbk.GBK_EPERM_MAP
    bad[3] = g.n + 5
    with pytest.raises(gbk.GbkError):
        ctx.relabel(dg, bad)


# ------------------------------------------------------------------ PageRank
# Tolerances: SpMV (Jacobi) is order-independent apart from float summation
# order, so scores match the oracle's to 1e-4 relative at the same iteration
# count.  Gauss-Seidel depends on the schedule in the reference itself, so
# it is held to the reference's own verifier and to the FP64 fixed point.
@pytest.mark.parametrize("scale", [10, 14])
def test_pr_spmv_matches_oracle(ctx, scale):
    g = O.kron(scale)
    dg = ctx.upload_graph(g)
    for iters in (1, 3, 20):
        s, st = ctx.pr(dg, gbk.PR_SPMV, max_iters=iters, tol=0.0, fixed=True)
        ref, it, err = O.pr_spmv(g, max_iters=iters, tol=0.0, fixed=True)
        assert st.iterations == it == iters
        np.testing.assert_allclose(s, ref, rtol=1e-4, atol=1e-9)
        # the L1 change is a sum of float differences: near convergence it
        # sits at the float rounding floor (~1e-8), where only an absolute
        # bound is meaningful
        assert st.final_error == pytest.approx(err, rel=1e-3, abs=1e-6)


def test_pr_spmv_convergence_iterations(ctx):
    g = O.kron(13)
    dg = ctx.upload_graph(g)
    s, st = ctx.pr(dg, gbk.PR_SPMV, max_iters=20, tol=1e-4)
    ref, it, err = O.pr_spmv(g, max_iters=20, tol=1e-4)
    assert st.iterations == it
    np.testing.assert_allclose(s, ref, rtol=1e-4, atol=1e-9)


@pytest.mark.parametrize("scale", [12, 15])
def test_pr_gs_verifies_and_reaches_fixed_point(ctx, scale):
    g = O.kron(scale)
    dg = ctx.upload_graph(g)
    s, st = ctx.pr(dg, gbk.PR_GS, max_iters=20, tol=1e-4)
    assert st.iterations <= 20 and st.final_error < 1e-4
    assert O.pr_residual(g, s) < 1e-4               # PRVerifier (pr.cc:117-134)
    exact = O.pr_fp64(g)
    l1 = float(np.abs(s - exact).sum())
    print(f"PR-GS scale {scale}: iters {st.iterations}, L1 vs FP64 {l1:.3e}")
    assert l1 < 1e-3
    s2, st2 = ctx.pr(dg, gbk.PR_GS, max_iters=200, tol=1e-9)
    np.testing.assert_allclose(s2, exact, rtol=1e-4, atol=1e-10)


def test_pr_heavy_rows_and_directed(ctx):
    g = _star(9000)
    dg = ctx.upload_graph(g)
    s, _ = ctx.pr(dg, gbk.PR_SPMV, max_iters=5, tol=0, fixed=True)
    ref, _, _ = O.pr_spmv(g, max_iters=5, tol=0, fixed=True)
    np.testing.assert_allclose(s, ref, rtol=1e-4)
    rng = np.random.default_rng(1)
    u = rng.integers(0, 500, 5000).astype(np.int32)
    v = (rng.zipf(1.3, 5000) % 500).astype(np.int32)
    gd = O.make_graph(u, v, symmetrize=False, n=500)
    ddg = ctx.upload_graph(gd)
    s, _ = ctx.pr(ddg, gbk.PR_SPMV, max_iters=8, tol=0, fixed=True)
    ref, _, _ = O.pr_spmv(gd, max_iters=8, tol=0, fixed=True)
    np.testing.assert_allclose(s, ref, rtol=1e-4, atol=1e-9)


# --------------------------------------------- Gauss-Seidel in row blocks
# The Gauss-Seidel sweep runs in pr_blocks tile-aligned row blocks, in order:
# a row of block b reads the new contributions of the rows before block b
# and the previous sweep's for the rest (one block: contributions updated in
# place, as the reference does; held to the verifier and the fixed point
# above).  Tiles start every 64 rows and
# wherever the prefix of light in-degrees (rows of at most 2048 in-edges)
# crosses a multiple of 2048; of T tiles, block b starts at tile b*T//B.
def _heavy_in_rows():
    """Directed graph whose in-rows sit at and around the heavy threshold and
    its chunk size, some far above it."""
    rng = np.random.default_rng(5)
    n = 20000
    sizes = [1, 64, 1023, 1024, 1025, 2047, 2048, 2049, 4096, 4097, 9000, 15000]
    us, vs = [], []
    for i, k in enumerate(sizes * 3):
        vs.append(np.full(k, (i * 613) % n, np.int32))
        us.append(rng.choice(n, k, replace=False).astype(np.int32))
    return O.make_graph(np.concatenate(us), np.concatenate(vs), symmetrize=False, n=n)


def _in_csr(g):
    return (g.in_off, g.in_nbr) if g.directed else (g.off, g.nbr)


def _block_starts(g, blocks, heavy=2048):
    in_off, _ = _in_csr(g)
    deg = np.diff(in_off)
    light = np.where(deg > heavy, 0, deg)
    lp = np.concatenate([[0], np.cumsum(light)[:-1]])
    v = np.arange(g.n)
    flag = (v % 64 == 0)
    flag[1:] |= (lp[1:] // 2048) != (lp[:-1] // 2048)
    tiles = np.append(np.flatnonzero(flag), g.n)
    T = len(tiles) - 1
    return np.array([int(tiles[T * b // blocks]) for b in range(blocks + 1)])


def _block_gs_fp64(g, blocks, iters):
    """The sweep above in float64 (pr.cc:40-94 with the block read rule)."""
    n = g.n
    in_off, in_nbr = _in_csr(g)
    out_deg = np.diff(g.off).astype(np.float64)
    rows = np.repeat(np.arange(n), np.diff(in_off))
    starts = _block_starts(g, blocks)
    s = np.full(n, 1.0 / n)
    c = np.where(out_deg > 0, s / np.maximum(out_deg, 1), 0.0)
    for _ in range(iters):
        new = c.copy()
        for lo, hi in zip(starts[:-1], starts[1:]):
            if hi <= lo:
                continue
            e0, e1 = in_off[lo], in_off[hi]
            src = in_nbr[e0:e1]
            val = np.where(src < lo, new[src], c[src])
            tot = np.bincount(rows[e0:e1] - lo, weights=val, minlength=hi - lo)
            s[lo:hi] = 0.15 / n + 0.85 * tot
            new[lo:hi] = np.where(out_deg[lo:hi] > 0, s[lo:hi] / np.maximum(out_deg[lo:hi], 1), 0.0)
        c = new
    return s


def _jacobi_fp64(g, iters):
    """PageRank-Jacobi (pr_spmv.cc:40-83) in float64."""
    return _block_gs_fp64(g, 1, iters)


@pytest.fixture
def pr_blocks(ctx):
    def use(count):
        ctx.set_param("pr_blocks", count)
    yield use
    ctx.set_param("pr_blocks", 0)


BLOCK_GRAPHS = {
    "kron13": lambda: O.kron(13),
    "kron12_sorted": lambda: O.relabel(O.kron(12), O.mapping_sort(O.kron(12))),
    "star9000": lambda: _star(9000),
    "heavy_in": _heavy_in_rows,
}


@pytest.mark.parametrize("name", list(BLOCK_GRAPHS))
@pytest.mark.parametrize("blocks", [2, 3, 16, 64])
def test_pr_gs_block_sweep_matches_fp64(ctx, pr_blocks, name, blocks):
    """One and three sweeps equal the float64 block sweep to 1e-5 (a row
    reading the wrong buffer is off by far more after one sweep)."""
    g = BLOCK_GRAPHS[name]()
    dg = ctx.upload_graph(g)
    pr_blocks(blocks)
    for iters in (1, 3):
        s, st = ctx.pr(dg, gbk.PR_GS, max_iters=iters, tol=0.0, fixed=True)
        assert st.iterations == iters
        np.testing.assert_allclose(s, _block_gs_fp64(g, blocks, iters), rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize("unroll", [1, 2, 4, 8])
def test_pr_gs_block_sweep_unrolls_match_fp64(ctx, pr_blocks, unroll):
    """Every chunks-in-flight setting of the block sweep gives the float64
    block sweep (the setting only changes the order of a row's float adds)."""
    g = BLOCK_GRAPHS["kron13"]()
    dg = ctx.upload_graph(g)
    pr_blocks(3)
    ctx.set_param("pr_unroll", unroll)
    try:
        s, st = ctx.pr(dg, gbk.PR_GS, max_iters=3, tol=0.0, fixed=True)
    finally:
        ctx.set_param("pr_unroll", 0)
    assert st.iterations == 3
    np.testing.assert_allclose(s, _block_gs_fp64(g, 3, 3), rtol=1e-5, atol=1e-12)


def test_block_gs_reference_model_pins():
    """The float64 model itself: one block is the oracle's Jacobi, and more
    blocks are not (so the tests above tell the read rules apart)."""
    g = O.kron(12)
    ref, _, _ = O.pr_spmv(g, max_iters=3, tol=0.0, fixed=True)
    np.testing.assert_allclose(_jacobi_fp64(g, 3), ref, rtol=1e-5)
    assert np.max(np.abs(_block_gs_fp64(g, 16, 1) / _jacobi_fp64(g, 1) - 1)) > 1e-2


@pytest.mark.parametrize("blocks", [4, 16, 64])
def test_pr_gs_blocks_converge(ctx, pr_blocks, blocks):
    """The fixed-order sweep converges in about the oracle's Gauss-Seidel
    sweeps (Kronecker scale 15: 13 in the float64 model for 4-256 blocks,
    the one-thread oracle 13; Gauss-Seidel is not faster than Jacobi at
    this size, as the reference's is not)."""
    g = O.kron(15)
    dg = ctx.upload_graph(g)
    pr_blocks(blocks)
    s, st = ctx.pr(dg, gbk.PR_GS, max_iters=20, tol=1e-4)
    _, it_ref, _ = O.pr_gs(g, max_iters=20, tol=1e-4)
    print(f"block GS x{blocks}: iters {st.iterations} (oracle GS {it_ref})")
    assert st.final_error < 1e-4 and st.iterations <= it_ref + 2
    assert O.pr_residual(g, s) < 1e-4               # PRVerifier (pr.cc:117-134)


def test_pr_heavy_rows_finished_in_launch(ctx):
    """A low heavy threshold (read when the graph is built) makes most rows
    heavy: each is finished by the workgroup adding its last chunk."""
    g = O.kron(13)
    ctx.set_param("heavy_degree", 64)
    try:
        dg = ctx.upload_graph(g)
    finally:
        ctx.set_param("heavy_degree", 2048)
    s, _ = ctx.pr(dg, gbk.PR_SPMV, max_iters=6, tol=0.0, fixed=True)
    ref, _, _ = O.pr_spmv(g, max_iters=6, tol=0.0, fixed=True)
    np.testing.assert_allclose(s, ref, rtol=1e-4, atol=1e-9)
    s, _ = ctx.pr(dg, gbk.PR_GS, max_iters=3, tol=0.0, fixed=True)
    # tiles of a 64-threshold graph differ from the default's: compare the
    # fixed point instead
    s2, _ = ctx.pr(dg, gbk.PR_GS, max_iters=200, tol=1e-9)
    np.testing.assert_allclose(s2, O.pr_fp64(g), rtol=1e-4, atol=1e-10)


# ------------------------------------------------- device mapping producers
# -o 0, 2, 5, 6, 7 made on the GPU (a stable sort by an out-degree key) are
# the host producers' mappings exactly (those are pinned to the oracle in
# tests/test_host.py); n spans one to many 2048-key chunks and merge passes.
def _odd_graph():
    rng = np.random.default_rng(9)
    n = 70001                                   # not a power of two: ragged last chunk
    u = rng.integers(0, n, 300000).astype(np.int32)
    v = (rng.zipf(1.6, 300000) % n).astype(np.int32)
    return O.make_graph(u, v, n=n)


MAP_GRAPHS = {
    "kron10": lambda: O.kron(10),
    "kron16": lambda: O.kron(16),
    "uniform14": lambda: O.kron(14, uniform=True),   # many degree ties
    "star5000": lambda: _star(5000),
    "odd70001": _odd_graph,
    "directed": lambda: _heavy_in_rows(),
}


@pytest.mark.parametrize("name", list(MAP_GRAPHS))
def test_device_mappings_equal_host(ctx, name):
    g = MAP_GRAPHS[name]()
    dg = ctx.upload_graph(g)
    for algo in (0, 1, 2, 5, 6, 7):
        ids, _ = ctx.mapping(dg, algo)
        assert np.array_equal(ids, H.mapping(algo, g.off)), (name, algo)


def test_relabel_by_equals_relabel(ctx):
    g = O.kron(13, weighted=True)
    dg = ctx.upload_graph(g)
    for algo in (1, 2, 5, 6):
        r1, _, ids, _ = ctx.relabel_by(dg, algo)
        assert np.array_equal(ids, H.mapping(algo, g.off))
        _assert_same_graph(r1, O.relabel(g, ids))
        r1.free()
    with pytest.raises(gbk.GbkError):
        ctx.mapping(dg, 3)                      # HubSort stays on the host


def _path(n):
    u = np.arange(max(n - 1, 1), dtype=np.int32) % max(n, 1)
    v = (u + 1) % max(n, 1)
    return O.make_graph(u[:max(n - 1, 0)], v[:max(n - 1, 0)], n=n)


@pytest.mark.parametrize("n", [2, 3, 1000, 70001, 1 << 22])
def test_device_random_mapping_equals_reference(ctx, n):
    """-o 1 on the device (parallel Fisher-Yates by deterministic
    reservations) equals the host's SplitMix64 Fisher-Yates, which equals the
    reference's GenerateRandomMapping (tests/test_oracle_ref.py)."""
    g = _path(n)
    dg = ctx.upload_graph(g)
    ids, ms = ctx.mapping(dg, 1)
    assert np.array_equal(ids, H.mapping_random(n)), n
    print(f"random mapping n={n}: {ms:.3f} ms on the device", flush=True)
    dg.free()


@pytest.mark.parametrize("blocks", [4, 16])
def test_pr_gs_inplace_blocks_reach_fixed_point(ctx, pr_blocks, blocks):
    """pr_inplace: ordered blocks that update contributions in place (rows may
    see same-sweep values of their own block, as in the reference's relaxed
    sweep): held to the verifier and the FP64 fixed point."""
    g = O.kron(14)
    dg = ctx.upload_graph(g)
    pr_blocks(blocks)
    ctx.set_param("pr_inplace", 1)
    try:
        s, st = ctx.pr(dg, gbk.PR_GS, max_iters=30, tol=1e-4)
        assert st.final_error < 1e-4 and O.pr_residual(g, s) < 1e-4
        s2, _ = ctx.pr(dg, gbk.PR_GS, max_iters=200, tol=1e-9)
        np.testing.assert_allclose(s2, O.pr_fp64(g), rtol=1e-4, atol=1e-10)
    finally:
        ctx.set_param("pr_inplace", 0)
