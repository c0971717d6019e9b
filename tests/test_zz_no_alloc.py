"""gbk_pr_noalloc (PageRank with its workspace in a fixed device array) against
gbk_pr, through the C ABI (GPU).

Two calls in a row, the second on a smaller graph (the array keeps the first
call's contents, so the second proves each call initialises what it reads):
Jacobi within 1e-4 relative of gbk_pr at the same iteration count, Gauss-
Seidel run to tol 1e-9 within 1e-4 relative of gbk_pr's, both against the
oracle too.  A workspace beyond the array's size is refused with EINVAL.
"""
import ctypes as C

import numpy as np
import pytest

# torch first: its bundled HIP runtime and /opt/rocm's share the soname
# libamdhip64.so.7, so whichever loads first serves the process, and torch
# only finds the GPU with its own.  libgbk.so is loaded lazily, later.
import torch

import oracle as O
from graphbrew_amd import gbk

pytestmark = pytest.mark.gpu


def _noalloc(ctx, dg, variant, max_iters, tol, fixed=False):
    scores = torch.empty(dg.n, dtype=torch.float32, device="cuda")
    report = torch.full((3,), -7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = gbk.lib().gbk_pr_noalloc(ctx.h, C.c_void_p(dg.handle), int(variant), int(max_iters), float(tol),
                                  int(fixed), C.c_void_p(scores.data_ptr()), C.c_void_p(report.data_ptr()))
    ctx._check(rc)
    ctx.sync()
    return scores.cpu().numpy(), report.cpu().numpy()


def test_noalloc_pagerank_equals_gbk_pr_two_calls(ctx):
    assert gbk.lib().gbk_pr_noalloc_max_bytes() == 64 << 20
    for scale in (13, 11):  # second call smaller than the first
        g = O.kron(scale)
        dg = ctx.upload_graph(g)
        s, rep = _noalloc(ctx, dg, gbk.PR_SPMV, 20, 1e-4)
        ref, st = ctx.pr(dg, gbk.PR_SPMV, max_iters=20, tol=1e-4)
        assert int(rep[0]) == st.iterations and int(rep[2]) == st.converged_iter
        np.testing.assert_allclose(rep[1], st.final_error, rtol=1e-6)
        np.testing.assert_allclose(s, ref, rtol=1e-4, atol=1e-10)
        es, it, _ = O.pr_spmv(g, max_iters=20, tol=1e-4)
        assert it == int(rep[0])
        np.testing.assert_allclose(s, es, rtol=1e-4, atol=1e-10)
        gs, grep = _noalloc(ctx, dg, gbk.PR_GS, 200, 1e-9)
        gref, gst = ctx.pr(dg, gbk.PR_GS, max_iters=200, tol=1e-9)
        assert grep[1] < 1e-9 and gst.final_error < 1e-9
        np.testing.assert_allclose(gs, gref, rtol=1e-4, atol=1e-10)
        np.testing.assert_allclose(gs, O.pr_fp64(g), rtol=1e-4, atol=1e-10)
        dg.free()


def test_noalloc_refuses_oversized_workspace(ctx):
    g = O.kron(10)
    dg = ctx.upload_graph(g)
    scores = torch.zeros(dg.n, dtype=torch.float32, device="cuda")
    # ten million iterations need 80 MB of per-iteration changes
    rc = gbk.lib().gbk_pr_noalloc(ctx.h, C.c_void_p(dg.handle), gbk.PR_SPMV, 10_000_000, 1e-4, 0,
                                  C.c_void_p(scores.data_ptr()), None)
    assert rc == gbk.GBK_EINVAL
    assert "workspace" in gbk.lib().gbk_last_error(ctx.h).decode()
    ctx.sync()
    assert float(scores.abs().sum().cpu()) == 0.0
    dg.free()
