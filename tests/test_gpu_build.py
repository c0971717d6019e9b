"""GPU CSR construction (gbk_graph_build) equals the oracle's MakeGraphFromEL
+ SquishGraph bit for bit."""
import os

import numpy as np
import pytest

import oracle as O
from graphbrew_amd import host as H
from tests import bruteforce as B

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "graphs")


def _same(dg, ref):
    d = dg.download()
    assert dg.n == ref.n
    assert np.array_equal(d["off"], ref.off) and np.array_equal(d["nbr"], ref.nbr)
    if ref.wgt is not None:
        assert np.array_equal(d["wgt"], ref.wgt)
    if ref.directed:
        assert np.array_equal(d["in_off"], ref.in_off) and np.array_equal(d["in_nbr"], ref.in_nbr)
        if ref.in_wgt is not None:
            assert np.array_equal(d["in_wgt"], ref.in_wgt)


@pytest.mark.parametrize("scale,uniform,weighted", [(12, False, False), (14, False, True),
                                                    (13, True, False), (16, False, False)])
def test_build_generated(ctx, scale, uniform, weighted):
    u, v = H.generate_el(scale, 16, uniform)
    w = H.insert_weights(len(u)) if weighted else None
    dg, ms = ctx.build(u, v, w)
    _same(dg, O.make_graph(u, v, w))


@pytest.mark.parametrize("name,weighted", [("tiny.el", False), ("isolated.el", False),
                                           ("weighted.wel", True), ("weighted_isolated.wel", True)])
@pytest.mark.parametrize("sym", [True, False])
@pytest.mark.parametrize("keep_self", [False, True])
def test_build_fixtures(ctx, name, weighted, sym, keep_self):
    u, v, w = B.read_el(os.path.join(GOLD, name), weighted)
    dg, _ = ctx.build(u, v, w, symmetrize=sym, keep_self=keep_self)
    _same(dg, O.make_graph(u, v, w, symmetrize=sym, keep_self=keep_self))


def test_build_duplicates_keep_smallest_weight(ctx):
    rng = np.random.default_rng(4)
    u = rng.integers(0, 300, 20000).astype(np.int32)
    v = (rng.zipf(1.3, 20000) % 300).astype(np.int32)
    w = rng.integers(1, 9, 20000).astype(np.int32)   # many ties
    for sym in (True, False):
        dg, _ = ctx.build(u, v, w, symmetrize=sym)
        _same(dg, O.make_graph(u, v, w, symmetrize=sym))
