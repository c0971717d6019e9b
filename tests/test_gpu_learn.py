"""Device features and the selector MLP's kernels (GPU).

Features are checked against the same statistics computed with numpy on the
host copy of the graph (degree moments, edge spans, sampled clustering with
the same sample rule), and degree statistics against their invariance under
relabelling.  The MLP's forward pass and gradient are checked against the
float64 oracle, which is itself pinned by finite differences in
tests/test_oracle.py; training must reduce the loss of a nonlinear fit.
"""
import numpy as np
import pytest

import oracle as O
from graphbrew_amd import gbk
from graphbrew_amd import host as H

pytestmark = pytest.mark.gpu


def _features_numpy(g):
    n, off, nbr = g.n, g.off.astype(np.int64), g.nbr.astype(np.int64)
    m = len(nbr)
    deg = np.diff(off).astype(np.float64)
    mean = deg.mean()
    avg = m // n
    src = np.repeat(np.arange(n), np.diff(off))
    span = np.abs(nbr - src)
    cc, cnt = 0.0, 0
    for w in range(1024):
        u = (w * 0x9E3779B97F4A7C15 % (1 << 64)) % n
        row = nbr[off[u]:off[u + 1]][:64]
        d = len(row)
        if d < 2:
            continue
        closed = 0
        for i in range(d):
            a = row[i]
            na = nbr[off[a]:off[a + 1]]
            closed += int(np.isin(row[i + 1:], na).sum())
        cc += closed / (d * (d - 1) / 2)
        cnt += 1
    ws = 8.0 * (n + 1) + 4.0 * m + 8.0 * n
    return np.array([
        np.log2(n), np.log2(max(m, 1)), mean, np.log2(1 + deg.max()),
        np.sqrt(max(0.0, (deg ** 2).mean() - mean ** 2)) / mean,
        deg[deg > avg].sum() / max(m, 1), cc / cnt if cnt else 0.0,
        np.log2(1 + span).sum() / max(m, 1) / max(1.0, np.log2(n)),
        (span < 64).mean(), (span < 4096).mean(),
        np.log2(ws / (4 << 20)), np.log2(ws / (256 << 20))])


@pytest.mark.parametrize("make", [lambda: O.kron(11), lambda: O.kron(10, uniform=True),
                                  lambda: O.make_graph(np.arange(999, dtype=np.int32),
                                                       np.arange(1, 1000, dtype=np.int32))])
def test_features_match_numpy(ctx, make):
    g = make()
    dg = ctx.upload_graph(g)
    f = ctx.features(dg)
    ref = _features_numpy(g)
    assert len(f) == gbk.NUM_FEATURES
    np.testing.assert_allclose(f, ref, rtol=1e-5, atol=1e-6)
    dg.free()


@pytest.mark.parametrize("make", [lambda: O.kron(12),
                                  lambda: O.make_graph(np.arange(19999, dtype=np.int32),
                                                       np.arange(1, 20000, dtype=np.int32))])
def test_degree_features_invariant_under_relabel(ctx, make):
    g = make()
    dg = ctx.upload_graph(g)
    rg, _ = ctx.relabel(dg, H.mapping_random(g.n))
    a, b = ctx.features(dg), ctx.features(rg)
    np.testing.assert_allclose(a[:6], b[:6], rtol=1e-6)
    np.testing.assert_allclose(a[10:], b[10:], rtol=1e-6)
    if g.n == 20000:
        # the path's own order is local; a random one is not (the generator
        # already permutes Kronecker ids, so there identity is random too)
        assert a[8] == 1.0 and b[8] < 0.02 and b[7] > 5 * a[7]
    dg.free()
    rg.free()


def _data(rows, d, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (rows, d)).astype(np.float32)
    y = (X[:, 0] * X[:, 1] + np.sin(2 * X[:, 2]) + 0.5 * X[:, 3 % d]).astype(np.float32)
    return X, y


@pytest.mark.parametrize("d,rows", [(5, 1), (28, 37), (48, 300)])
def test_mlp_forward_and_gradient_match_oracle(ctx, d, rows):
    net = ctx.mlp(d, seed=d)
    p = net.params
    assert p.size == 64 * d + 64 + 64 * 64 + 64 + 64 + 1
    X, y = _data(rows, d, seed=rows)
    np.testing.assert_allclose(net.forward(X), O.mlp_forward(p, d, X), rtol=1e-4, atol=1e-5)
    loss, g = net.grad(X, y)
    rloss, rg = O.mlp_grad(p, d, X, y)
    assert abs(loss - rloss) <= 1e-4 * max(1.0, rloss)
    scale = np.abs(rg).max()
    assert np.abs(g - rg).max() <= 1e-4 * scale + 1e-6
    net.close()


def test_mlp_params_roundtrip_and_rejects_bad_shapes(ctx):
    net = ctx.mlp(7, seed=1)
    p = np.random.default_rng(0).normal(0, 0.1, net.num_params).astype(np.float32)
    net.params = p
    assert np.array_equal(net.params, p)
    with pytest.raises(ValueError):
        net.forward(np.zeros((3, 8), np.float32))
    with pytest.raises(gbk.GbkError):
        ctx.mlp(49)
    net.close()


def test_mlp_training_fits_nonlinear_target(ctx):
    d = 6
    X, y = _data(2048, d, seed=5)
    net = ctx.mlp(d, seed=2)
    before = float(((net.forward(X) - y) ** 2).mean())
    curve = net.train(X, y, epochs=60, batch=128, lr=3e-3, seed=1)
    after = float(((net.forward(X) - y) ** 2).mean())
    assert curve[-1] < curve[0]
    assert after < 0.1 * before
    # better than the best linear fit, which cannot represent x0 * x1
    A = np.c_[X, np.ones(len(X))]
    lin = np.linalg.lstsq(A, y, rcond=None)[0]
    assert after < 0.5 * float(((A @ lin - y) ** 2).mean())
    net.close()
