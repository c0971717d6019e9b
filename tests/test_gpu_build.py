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


@pytest.mark.parametrize("weighted", [False, True])
def test_build_thousands_of_copies(ctx, weighted):
    """Raw rows holding thousands of copies of one edge (sorted as buckets of
    one repeated key) squish to the reference's CSR."""
    rng = np.random.default_rng(9)
    bu = rng.integers(0, 5000, 30000)
    bv = rng.integers(0, 5000, 30000)
    u = np.concatenate([np.zeros(6000), np.full(2500, 7), bu]).astype(np.int32)
    v = np.concatenate([np.ones(6000), np.full(2500, 4999), bv]).astype(np.int32)
    w = rng.integers(1, 4, len(u)).astype(np.int32) if weighted else None
    for sym in (True, False):
        dg, _ = ctx.build(u, v, w, symmetrize=sym)
        _same(dg, O.make_graph(u, v, w, symmetrize=sym))


@pytest.mark.parametrize("case", ["bad_id", "negative_id", "decreasing", "bad_in_id"])
def test_upload_rejects_malformed_csr(ctx, case):
    """gbk_graph_upload checks a caller's CSR on the host before any kernel
    sees it: out-of-range neighbour ids and decreasing offsets are EINVAL."""
    from graphbrew_amd import gbk
    off = np.array([0, 2, 3, 4], np.int64)
    nbr = np.array([1, 2, 0, 0], np.int32)
    in_off, in_nbr, directed = None, None, False
    if case == "bad_id":
        nbr[2] = 3
    elif case == "negative_id":
        nbr[0] = -1
    elif case == "decreasing":
        off = np.array([0, 3, 2, 4], np.int64)
    else:
        directed = True
        in_off = np.array([0, 2, 3, 4], np.int64)
        in_nbr = np.array([1, 2, 0, 7], np.int32)
    with pytest.raises(gbk.GbkError) as e:
        ctx.upload(3, off, nbr, in_off=in_off, in_nbr=in_nbr, directed=directed)
    assert e.value.code == gbk.GBK_EINVAL
    ok = ctx.upload(3, np.array([0, 2, 3, 4], np.int64), np.array([1, 2, 0, 0], np.int32))
    ok.free()
