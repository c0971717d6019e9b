"""Many small graphs in one launch (GPU): the union graph equals its parts
shifted, CC labels of every part are bit-exact to the oracle run on that part
alone, and PR of every part matches a run on that part alone (SpMV: 1e-4
relative to the oracle at the part's own iteration count, which equals the
one-graph run's; Gauss-Seidel: the reference's verifier and the FP64 fixed
point, as for one graph)."""
import os

import numpy as np
import pytest

import oracle as O
from graphbrew_amd import gbk
from tests import bruteforce as B

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "graphs")


def _parts():
    rng = np.random.default_rng(7)
    parts = [O.kron(s) for s in (6, 9, 11, 8, 12, 7)]
    parts += [O.kron(s, uniform=True) for s in (8, 10)]
    u, v, _ = B.read_el(os.path.join(GOLD, "disconnected.el"), False)
    parts.append(O.make_graph(u, v))
    u, v, _ = B.read_el(os.path.join(GOLD, "path.el"), False)
    parts.append(O.make_graph(u, v))
    star = O.make_graph(np.zeros(5000, np.int32), np.arange(1, 5001, dtype=np.int32))
    parts.append(star)  # one heavy row
    for _ in range(20):
        n = int(rng.integers(2, 300))
        m = int(rng.integers(1, 4 * n))
        parts.append(O.make_graph(rng.integers(0, n, m).astype(np.int32),
                                  rng.integers(0, n, m).astype(np.int32), n=n))
    return parts


@pytest.fixture(scope="module")
def batch(ctx):
    parts = _parts()
    dgs = [ctx.upload_graph(g) for g in parts]
    ug, vbase = ctx.graph_union(dgs)
    yield parts, dgs, ug, vbase
    for d in dgs:
        d.free()
    ug.free()


def test_union_is_the_parts_shifted(ctx, batch):
    parts, _, ug, vbase = batch
    assert vbase[0] == 0 and vbase[-1] == ug.n == sum(g.n for g in parts)
    assert ug.m == sum(g.m for g in parts)
    d = ug.download()
    for j, g in enumerate(parts):
        lo, hi = vbase[j], vbase[j + 1]
        off = d["off"][lo:hi + 1] - d["off"][lo]
        assert np.array_equal(off, g.off)
        assert np.array_equal(d["nbr"][d["off"][lo]:d["off"][hi]] - lo, g.nbr)


@pytest.mark.parametrize("variant", [gbk.CC_AFFOREST, gbk.CC_SV])
def test_cc_many_matches_each_part(ctx, batch, variant):
    parts, _, ug, vbase = batch
    comp, _ = ctx.cc_many(ug, vbase, variant)
    for j, g in enumerate(parts):
        ref = O.cc_afforest(g)
        assert np.array_equal(comp[vbase[j]:vbase[j + 1]], ref), j


def test_pr_many_spmv_matches_each_part(ctx, batch):
    parts, dgs, ug, vbase = batch
    s, iters, st = ctx.pr_many(ug, vbase, gbk.PR_SPMV, max_iters=20, tol=1e-4)
    assert st.iterations == iters.max()
    for j, g in enumerate(parts):
        one, st1 = ctx.pr(dgs[j], gbk.PR_SPMV, max_iters=20, tol=1e-4)
        # the same stopping rule per part; float summation order may move a
        # change that sits exactly at the tolerance by one iteration
        assert abs(int(iters[j]) - st1.iterations) <= 1, j
        ref, it, _ = O.pr_spmv(g, max_iters=int(iters[j]), tol=0.0, fixed=True)
        np.testing.assert_allclose(s[vbase[j]:vbase[j + 1]], ref, rtol=1e-4, atol=1e-9)


def test_pr_many_fixed_runs_every_part_to_the_end(ctx, batch):
    parts, _, ug, vbase = batch
    s, iters, _ = ctx.pr_many(ug, vbase, gbk.PR_SPMV, max_iters=5, tol=1e-4, fixed=True)
    assert np.all(iters == 5)
    for j, g in enumerate(parts):
        ref, _, _ = O.pr_spmv(g, max_iters=5, tol=0.0, fixed=True)
        np.testing.assert_allclose(s[vbase[j]:vbase[j + 1]], ref, rtol=1e-4, atol=1e-9)


def test_pr_many_gauss_seidel_reaches_each_fixed_point(ctx, batch):
    parts, _, ug, vbase = batch
    s, iters, _ = ctx.pr_many(ug, vbase, gbk.PR_GS, max_iters=200, tol=1e-9)
    for j, g in enumerate(parts):
        part = s[vbase[j]:vbase[j + 1]]
        np.testing.assert_allclose(part, O.pr_fp64(g), rtol=1e-4, atol=1e-10)


def test_many_rejects_bad_bases(ctx, batch):
    _, _, ug, vbase = batch
    bad = vbase.copy()
    bad[-1] -= 1
    with pytest.raises(gbk.GbkError):
        ctx.cc_many(ug, bad)
    with pytest.raises(gbk.GbkError):
        ctx.pr_many(ug, np.array([0, 0, ug.n], np.int64))
