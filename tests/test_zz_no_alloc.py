"""The allocation-free weighted-PageRank forward (suminc_wpr_forward_noalloc)
against the allocating suminc_wpr_forward, through the C ABI: two calls in a row
on one context, the second on a smaller graph, each checked against the
allocating forward on the same inputs (1e-6 of the largest rank: the two differ
only in the order of the per-chunk float atomics) and against the FP64 oracle
(L1 <= 1e-6).  Size and state errors come back as codes."""
import ctypes as C

import numpy as np
import pytest

import oracle
from suminc_amd import _lib as L
from tests.conftest import DAMPING

pytestmark = pytest.mark.gpu


def _noalloc(ctx, pd, rd):
    s = L.Stats()
    rc = ctx._lib.suminc_wpr_forward_noalloc(ctx._h, C.c_void_p(pd.data_ptr()), C.c_void_p(rd.data_ptr()),
                                             C.byref(s))
    return rc, s.as_dict()


def _personalization(n, rng, k=64):
    p = np.zeros(n, np.float32)
    idx = rng.choice(n, k, replace=False)
    p[idx] = rng.dirichlet(np.ones(k)).astype(np.float32)
    return p


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_noalloc_forward_matches_forward(gpu, prec):
    pp = gpu.PR_FP32_KAHAN if prec == "fp32" else gpu.PR_FP64
    rng = np.random.default_rng(7)
    with gpu.Context(pr_precision=pp, termcheck_threshold=1e-10) as ctx:
        for scale, seed in ((14, 4), (11, 6)):  # the second call on a smaller graph
            g = oracle.Graph.rmat(scale, 16, seed=seed, weights="f32")
            off, col, w = g.csr()
            ctx.upload_csr(off, col, w)
            p = _personalization(g.n, rng)
            ctx.wpr_set_weights(w)
            r_ref, _ = ctx.wpr_forward(p)
            pd = ctx.device_array(p)
            rd = ctx.device_array(np.full(g.n, np.nan, np.float32))
            rc, st = _noalloc(ctx, pd, rd)
            assert rc == 0, ctx._lib.suminc_last_error(ctx._h)
            r = rd.download()
            pd.free()
            rd.free()
            assert st["rounds"] > 1 and st["residual"] < 1e-10
            assert np.abs(r - r_ref).max() <= 1e-6 * np.abs(r_ref).max(), (scale, np.abs(r - r_ref).max())
            r64, _ = g.wpr_forward(w, p.astype(np.float64), DAMPING)
            assert np.abs(r - r64).sum() <= 1e-6, (scale, np.abs(r - r64).sum())


def test_noalloc_forward_errors(gpu):
    rng = np.random.default_rng(8)
    with gpu.Context(pr_precision=gpu.PR_FP64, termcheck_threshold=1e-10) as ctx:
        g = oracle.Graph.rmat(10, 16, seed=3, weights="f32")
        off, col, w = g.csr()
        ctx.upload_csr(off, col, w)
        p = _personalization(g.n, rng, 8)
        pd = ctx.device_array(p)
        rd = ctx.device_array(np.zeros(g.n, np.float32))
        ctx.wpr_set_weights(w)
        rc, _ = _noalloc(ctx, pd, rd)  # coefficients not prepared yet
        assert rc == gpu.E_STATE
        ctx.wpr_forward(p)
        s = L.Stats()
        rh = np.zeros(g.n, np.float32)
        rc = ctx._lib.suminc_wpr_forward_noalloc(ctx._h, C.c_void_p(pd.data_ptr()), L.ptr(rh), C.byref(s))
        assert rc == gpu.E_INVALID  # host memory for the rank
        assert _noalloc(ctx, pd, rd)[0] == 0
        pd.free()
        rd.free()
        # 4.19M vertices in FP64 need more than the fixed scratch array holds
        ctx.generate_rmat(22, 16, 0.57, 0.19, 0.19, seed=4, weights="f32")
        n = ctx.graph_info()[0]
        p = _personalization(n, rng, 8)
        ctx.wpr_forward(p)
        pd = ctx.device_array(p)
        rd = ctx.device_array(np.zeros(n, np.float32))
        rc, _ = _noalloc(ctx, pd, rd)
        assert rc == gpu.E_TOO_LARGE
        assert b"scratch" in ctx._lib.suminc_last_error(ctx._h)
        pd.free()
        rd.free()
