"""GPU parity: relabel (bit-exact) and PageRank (tolerances stated below)
against the CPU oracle, through the C ABI."""
import numpy as np
import pytest

import oracle as O
from graphbrew_amd import gbk

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


GRAPHS = {
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
