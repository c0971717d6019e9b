"""Parity at Kronecker scale 24 (16.8M vertices, ~520M directed edges), the
scale the path's targets are stated at (GBK_LARGE_SCALE overrides the
scale; about half a minute of host work for the oracle).  Size-independent checks are
used where the oracle would take too long: relabel under a random mapping
is checked row by row on a sample against the oracle's permutation rule,
every kernel's result against the oracle for a few sources."""
import os

import numpy as np
import pytest

import oracle as O
from graphbrew_amd import gbk
from graphbrew_amd import host as H

pytestmark = pytest.mark.gpu

SCALE = int(os.environ.get("GBK_LARGE_SCALE", "24"))
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def big(ctx):
    u, v = H.generate_el(SCALE, 16)
    w = H.insert_weights(len(u))
    g, _ = ctx.build(u, v, None, -1, True)
    gw, _ = ctx.build(u, v, w, g.n, True)
    del u, v, w
    d = g.download()
    host = O.Graph(n=g.n, off=d["off"], nbr=d["nbr"])
    dw = gw.download()
    host_w = O.Graph(n=gw.n, off=dw["off"], nbr=dw["nbr"], wgt=dw["wgt"])
    print(f"scale {SCALE}: n={g.n} m={g.m}", flush=True)
    yield g, gw, host, host_w
    g.free()
    gw.free()


def test_relabel_random_mapping(ctx, big):
    g, _, host, _ = big
    ids = H.mapping_random(g.n)
    rg, st = ctx.relabel(g, ids)
    d = rg.download()
    inv = np.empty_like(ids)
    inv[ids] = np.arange(g.n, dtype=np.int32)
    rng = np.random.default_rng(1)
    for r in rng.integers(0, g.n, 20000):
        u = inv[r]
        want = np.sort(ids[host.nbr[host.off[u]:host.off[u + 1]]])
        got = d["nbr"][d["off"][r]:d["off"][r + 1]]
        assert np.array_equal(got, want), r
    assert np.array_equal(d["org_ids"], inv)
    print(f"relabel: {st.device_ms:.1f} ms", flush=True)
    rg.free()


def test_bfs_cc_pr(ctx, big):
    g, _, host, _ = big
    srcs = O.source_picker(host, 3)[1]
    depth, _, _, _ = ctx.bfs_batch(g, srcs)
    for j, s in enumerate(srcs):
        assert np.array_equal(depth[j], O.bfs_depths(host, int(s)))
        parent, _ = ctx.bfs(g, int(s))
        assert O.bfs_verify(host, int(s), parent)
    comp, _ = ctx.cc(g, gbk.CC_AFFOREST)
    assert np.array_equal(comp, O.cc_afforest(host, threads=THREADS))
    comp_sv, _ = ctx.cc(g, gbk.CC_SV)
    assert np.array_equal(comp_sv, comp)
    s, st = ctx.pr(g, gbk.PR_SPMV, max_iters=3, tol=0.0, fixed=True)
    ref, _, _ = O.pr_spmv(host, max_iters=3, tol=0.0, fixed=True, threads=THREADS)
    np.testing.assert_allclose(s, ref, rtol=1e-4, atol=1e-10)


def test_sssp_and_bc(ctx, big):
    _, gw, _, host_w = big
    srcs = O.source_picker(host_w, 2)[1]
    dist, _, _, _ = ctx.sssp_batch(gw, srcs, 1024)
    for j, s in enumerate(srcs):
        ref = O.sssp_delta(host_w, int(s), 64, threads=THREADS)
        assert np.array_equal(dist[j], ref)
        one, _ = ctx.sssp(gw, int(s), 64)
        assert np.array_equal(one, ref)
    host = O.Graph(n=host_w.n, off=host_w.off, nbr=host_w.nbr)
    sc, _ = ctx.bc(gw, [int(srcs[0])])
    ref = O.bc(host, [int(srcs[0])], threads=THREADS)
    np.testing.assert_allclose(sc, ref, rtol=1e-4, atol=1e-6)
