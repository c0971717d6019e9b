"""PageRank's segmented sweep (GPU): the in-edges walked segment by segment
of the source range, each segment's contributions one L2-sized slice.

By default it runs on graphs above 2^19 vertices (the large and
configuration tests cover it there); here the "pr_segment" parameter forces
narrow segments on small graphs, so that every case of the layout is met:
heavy rows cut into pieces at the segment bounds (heavy_degree 64), directed
graphs, rows whose sources are not ascending (an uploaded CSR need not be
sorted), the most segments (32) and Gauss-Seidel blocks that publish their
contributions.  Bars: Jacobi (SpMV) within 1e-4 of the oracle's float
iteration and of the row-tile sweep; Gauss-Seidel within 1e-5 of the
row-tile sweep over the same fixed sweeps (both read the same
contributions; only the order of the float sums differs), and to
convergence in the same number of sweeps within 1e-4 of the oracle's
fixed point."""
import contextlib

import numpy as np
import pytest

import oracle as O
from graphbrew_amd import gbk

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def params(ctx, **kw):
    saved = {k: ctx.get_param(k) for k in kw}
    try:
        for k, v in kw.items():
            ctx.set_param(k, v)
        yield
    finally:
        for k, v in saved.items():
            ctx.set_param(k, v)


def _directed(n=12000, m=150000, seed=5):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m).astype(np.int32)
    v = (rng.integers(0, n, m) * rng.integers(0, n, m) // n).astype(np.int32)  # skewed heads
    return O.make_graph(u, v, symmetrize=False, n=n)


GRAPHS = {
    "kron14": lambda: O.kron(14),
    "uniform14": lambda: O.kron(14, uniform=True),
    "directed12000": _directed,
}


def _run(ctx, dg, variant, seg, iters, tol=0.0, fixed=True):
    with params(ctx, pr_segment=seg):
        return ctx.pr(dg, variant, max_iters=iters, tol=tol, fixed=fixed)


@pytest.mark.parametrize("heavy", [2048, 64])
@pytest.mark.parametrize("width", [9, 11])
@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_segmented_spmv(ctx, name, width, heavy):
    g = GRAPHS[name]()
    with params(ctx, heavy_degree=heavy):
        dg = ctx.upload_graph(g)
    try:
        exp, _, _ = O.pr_spmv(g, max_iters=6, tol=0.0, fixed=True)
        got, st = _run(ctx, dg, gbk.PR_SPMV, width, 6)
        row, _ = _run(ctx, dg, gbk.PR_SPMV, -1, 6)
        assert st.iterations == 6
        np.testing.assert_allclose(got, exp, rtol=1e-4, atol=1e-10)
        np.testing.assert_allclose(got, row, rtol=1e-4, atol=1e-10)
    finally:
        dg.free()


@pytest.mark.parametrize("heavy", [2048, 64])
@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_segmented_gauss_seidel_blocks(ctx, name, heavy):
    g = GRAPHS[name]()
    with params(ctx, heavy_degree=heavy):
        dg = ctx.upload_graph(g)
    try:
        with params(ctx, pr_blocks=4):
            got, _ = _run(ctx, dg, gbk.PR_GS, 10, 5)
            row, _ = _run(ctx, dg, gbk.PR_GS, -1, 5)
            np.testing.assert_allclose(got, row, rtol=1e-5, atol=1e-11)
            conv, st = _run(ctx, dg, gbk.PR_GS, 10, 50, tol=1e-6, fixed=False)
            conv_row, st_row = _run(ctx, dg, gbk.PR_GS, -1, 50, tol=1e-6, fixed=False)
        assert st.converged_iter > 0 and abs(st.converged_iter - st_row.converged_iter) <= 1
        exact = O.pr_fp64(g)
        np.testing.assert_allclose(conv, exact, rtol=1e-4, atol=1e-9)
        np.testing.assert_allclose(conv, conv_row, rtol=1e-4, atol=1e-10)
    finally:
        dg.free()


def test_segmented_unsorted_rows(ctx):
    """Rows whose sources are shuffled: light tiles take the lane-by-lane
    scatter, heavy rows one piece; the sums are the sorted graph's."""
    g = O.kron(13)
    rng = np.random.default_rng(3)
    nbr = g.nbr.copy()
    for u in range(g.n):
        a, b = g.off[u], g.off[u + 1]
        if b - a > 1:
            nbr[a:b] = rng.permutation(nbr[a:b])
    shuffled = O.Graph(n=g.n, off=g.off, nbr=nbr)
    assert not np.array_equal(nbr, g.nbr)
    exp, _, _ = O.pr_spmv(g, max_iters=6, tol=0.0, fixed=True)
    for heavy in (2048, 64):
        with params(ctx, heavy_degree=heavy):
            dg = ctx.upload_graph(shuffled)
        try:
            got, _ = _run(ctx, dg, gbk.PR_SPMV, 9, 6)
            np.testing.assert_allclose(got, exp, rtol=1e-4, atol=1e-10)
        finally:
            dg.free()


def test_segmented_most_segments_and_knob(ctx):
    """2^6-wide segments are widened until at most 32 remain; the knob
    rejects widths outside 6..26 and reads back what was set."""
    g = O.kron(11)
    dg = ctx.upload_graph(g)
    try:
        exp, _, _ = O.pr_spmv(g, max_iters=4, tol=0.0, fixed=True)
        got, _ = _run(ctx, dg, gbk.PR_SPMV, 6, 4)  # 2048 vertices: 2^6 gives 32 segments
        np.testing.assert_allclose(got, exp, rtol=1e-4, atol=1e-10)
    finally:
        dg.free()
    for bad in (-2, 1, 5, 27):
        with pytest.raises(gbk.GbkError):
            ctx.set_param("pr_segment", bad)
    with params(ctx, pr_segment=12):
        assert ctx.get_param("pr_segment") == 12
    assert ctx.get_param("pr_segment") == 0


def test_layout_kept_with_the_graph(ctx):
    """The first call builds the layout and later calls reuse it: repeated
    calls give the same scores (to 1e-6: the block sweeps fix what each row
    reads, but heavy rows' chunk sums are added by float atomics in any
    order), and a call with another segment width rebuilds it, after which
    the first width's scores come back."""
    g = O.kron(14)
    with params(ctx, heavy_degree=64):
        dg = ctx.upload_graph(g)
    try:
        with params(ctx, pr_blocks=4):
            a, _ = _run(ctx, dg, gbk.PR_GS, 9, 6)
            b, _ = _run(ctx, dg, gbk.PR_GS, 9, 6)
            np.testing.assert_allclose(b, a, rtol=1e-6, atol=1e-12)
            c, _ = _run(ctx, dg, gbk.PR_GS, 11, 6)
            np.testing.assert_allclose(c, a, rtol=1e-5, atol=1e-11)
            d, _ = _run(ctx, dg, gbk.PR_GS, 9, 6)
            np.testing.assert_allclose(d, a, rtol=1e-6, atol=1e-12)
        exp, _, _ = O.pr_spmv(g, max_iters=6, tol=0.0, fixed=True)
        got, _ = _run(ctx, dg, gbk.PR_SPMV, 9, 6)
        np.testing.assert_allclose(got, exp, rtol=1e-4, atol=1e-10)
    finally:
        dg.free()
