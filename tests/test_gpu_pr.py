"""Delta-PageRank on the device vs the oracle's closed form / FP64 runs.

Tolerances (the metric: L1 <= 1e-6 vs the FP64 oracle):
  * FP64 path: ||r - r*||_1 <= residual / (1 - d) + 1e-12 (the stop-rule bound).
  * FP32 + Kahan path: ||r - r*||_1 <= residual / (1 - d) + 1e-6, and the
    FP64 mass invariant holds to 1e-6.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.conftest import DAMPING

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def mass_defect(off, r, delta, d):
    nd = np.diff(off) > 0
    return abs(r.sum() + delta.sum() - d * r[nd].sum() - (1 - d))


def test_golden_test_base(gpu):
    exp = json.loads((GOLD / "test_base_expected.json").read_text())
    e = np.loadtxt(GOLD / "test_base.txt", dtype=np.int64, comments="#")
    g = oracle.Graph.from_edges(32, e[:, 0], e[:, 1])
    off, col, _ = g.csr()
    for prec in (gpu.PR_FP64, gpu.PR_FP32_KAHAN):
        with gpu.Context(pr_precision=prec, termcheck_threshold=1e-12) as ctx:
            ctx.upload_csr(off, col)
            ctx.pr_init()
            st = ctx.pr_run()
            r, d = ctx.pr_download()
        err = np.abs(r - np.array(exp["pagerank"])).sum()
        # FP32 deltas round each push to 2^-24 relative: ~6e-8 x the pushed mass
        tol = st["residual"] / (1 - DAMPING) + (1e-13 if prec == gpu.PR_FP64 else 1e-7)
        assert err <= tol, (prec, err, tol)


# band_kib=16: the banded pull (source bands of 4096 / 2048 vertices at FP32 /
# FP64 on 2^16 vertices, over compact positions in id order when compact_kib=1);
# compact_kib=1 alone: the pull over compact source positions by out-degree, used
# by default above 32 MiB of contributions (8.4M vertices in FP32); tile_kib: the
# tiled pull (destination tiles summed in LDS over source windows of 1 or 4 KiB,
# blocks of the hub rows cut into several tiles that add atomically), the default
# above 64 MiB of contributions when the binned pull is off
@pytest.mark.parametrize("scale,band_kib,compact_kib,tile_kib", [(12, 0, 0, -1), (16, 0, 0, -1), (16, 16, 0, -1),
                                                                 (16, 0, 1, -1), (16, 16, 1, -1), (16, 0, 0, 1),
                                                                 (16, 0, 1, 1), (16, 0, 1, 4), (12, 0, 0, 1)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_pagerank_vs_fp64_oracle(gpu, scale, band_kib, compact_kib, tile_kib, prec):
    g = oracle.Graph.rmat(scale, 16, seed=1)
    off, col, _ = g.csr()
    rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
    T = 1e-9 * g.n  # per-vertex threshold 1e-9 (the C1 setting)
    p = gpu.PR_FP64 if prec == "fp64" else gpu.PR_FP32_KAHAN
    with gpu.Context(pr_precision=p, termcheck_threshold=T, pull_band_kib=band_kib,
                     pull_compact_kib=compact_kib, pull_tile_kib=tile_kib) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        st = ctx.pr_run()
        r, d = ctx.pr_download()
    assert st["residual"] < T
    err = np.abs(r - rs).sum()
    extra = 1e-12 if prec == "fp64" else 1e-6
    assert err <= st["residual"] / (1 - DAMPING) + extra, (err, st)
    assert mass_defect(off, r, d, DAMPING) < (1e-12 if prec == "fp64" else 1e-6)
    assert st["edges_traversed"] > 0 and st["rounds"] >= 2
    # the reference-style CPU run with the same stop rule lands in the same band
    ro, do, sto = g.pr_ingress(DAMPING, T, max_rounds=100000)
    assert np.abs(ro - rs).sum() <= sto["residual"] / (1 - DAMPING) + 1e-12


# the binned pull: 1 or 4 KiB source windows (the value stream, then the summing
# pass by destination block; blocks cut into parts that add atomically when long),
# the first `direct` windows gathered by the summing pass itself (-1: the default,
# one; 0: every window through the stream; 1000: none through it), the default
# above 8 MiB of contributions
@pytest.mark.parametrize("scale,compact_kib,bin_kib,direct", [(16, 1, 1, -1), (16, 0, 4, -1), (16, 1, 1, 0),
                                                             (16, 1, 1, 1000), (12, 0, 1, 3), (16, 1, 4, 2)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_binned_pull_vs_fp64_oracle(gpu, scale, compact_kib, bin_kib, direct, prec):
    g = oracle.Graph.rmat(scale, 16, seed=1)
    off, col, _ = g.csr()
    rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
    T = 1e-9 * g.n
    p = gpu.PR_FP64 if prec == "fp64" else gpu.PR_FP32_KAHAN
    with gpu.Context(pr_precision=p, termcheck_threshold=T, pull_compact_kib=compact_kib, pull_bin_kib=bin_kib,
                     pull_bin_direct=direct) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        st = ctx.pr_run()
        r, d = ctx.pr_download()
    assert st["residual"] < T
    err = np.abs(r - rs).sum()
    extra = 1e-12 if prec == "fp64" else 1e-6
    assert err <= st["residual"] / (1 - DAMPING) + extra, (err, st)
    assert mass_defect(off, r, d, DAMPING) < (1e-12 if prec == "fp64" else 1e-6)


def test_l1_target_fp32_kahan(gpu):
    """FP32 deltas + Kahan ranks reach L1 <= 1e-6 vs the FP64 fixed point."""
    g = oracle.Graph.rmat(16, 16, seed=3)
    off, col, _ = g.csr()
    rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
    with gpu.Context(pr_precision=gpu.PR_FP32_KAHAN, termcheck_threshold=1e-8) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        ctx.pr_run()
        r, _ = ctx.pr_download()
    assert np.abs(r - rs).sum() <= 1e-6


def test_round_cap_reports_not_converged(gpu):
    g = oracle.Graph.rmat(10, 16, seed=1)
    off, col, _ = g.csr()
    with gpu.Context(termcheck_threshold=1e-12, max_rounds=3) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        with pytest.raises(gpu.SumincError) as e:
            ctx.pr_run()
        assert e.value.code == -6
        st = ctx.pr_run(allow_not_converged=True)  # the state is still valid and can continue
        assert st["rounds"] == 3


def test_step_and_upload(gpu):
    g = oracle.Graph.rmat(11, 16, seed=1)
    off, col, _ = g.csr()
    r0 = np.zeros(g.n)
    d0 = np.full(g.n, (1 - DAMPING) / g.n)
    with gpu.Context(pr_precision=gpu.PR_FP64, termcheck_threshold=0.0) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_upload(r0, d0)
        res = [ctx.pr_step() for _ in range(3)]
        r, d = ctx.pr_download()
    assert res[0] == pytest.approx(1 - DAMPING, rel=1e-12)
    assert res[0] > res[1] > res[2]
    # three synchronous rounds with no threshold == three oracle Jacobi rounds
    ro, do, _ = g.pr_solve(DAMPING, tol=0.0, max_rounds=3)
    assert np.abs(r - ro).max() < 1e-15 and np.abs(d - do).max() < 1e-15


@pytest.mark.parametrize("band_kib,compact_kib,tile_kib,bin_kib", [(0, 0, -1, 0), (16, 0, -1, 0), (0, 1, -1, 0),
                                                                  (16, 1, -1, 0), (0, 0, 1, 0), (0, 1, 1, 0),
                                                                  (0, 1, -1, 1), (0, 0, -1, 2)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_incremental_reseed_matches_scratch(gpu, prec, band_kib, compact_kib, tile_kib, bin_kib):
    g = oracle.Graph.rmat(14, 16, seed=3)
    off, col, _ = g.csr()
    T = 1e-10 * g.n
    p = gpu.PR_FP64 if prec == "fp64" else gpu.PR_FP32_KAHAN
    with gpu.Context(pr_precision=p, termcheck_threshold=T, pull_band_kib=band_kib,
                     pull_compact_kib=compact_kib, pull_tile_kib=tile_kib, pull_bin_kib=bin_kib) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        ctx.pr_run()
        for b in range(3):
            n_ops = int(0.01 * g.m)
            ops = g.gen_batch(seed=3, batch_id=b, n_del=n_ops // 2, n_ins=n_ops - n_ops // 2)
            ctx.apply_batch(ops)
            g, _ = g.apply_batch(ops)
            st = ctx.pr_run()
            r, d = ctx.pr_download()
            rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
            err = np.abs(r - rs).sum()
            assert err <= st["residual"] / (1 - DAMPING) + (1e-11 if prec == "fp64" else 2e-6), (b, err, st)


def signed_mass_defect(off, r, delta, d):
    nd = np.diff(off) > 0
    return r.sum() + delta.sum() - d * r[nd].sum() - (1 - d)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_device_mass_balance_matches_host(gpu, prec):
    """suminc_stats.mass_defect (FP64 on the device) equals the same sum over the
    downloaded state, after a run and after an incremental batch."""
    g = oracle.Graph.rmat(14, 16, seed=5)
    off, col, _ = g.csr()
    p = gpu.PR_FP64 if prec == "fp64" else gpu.PR_FP32_KAHAN
    with gpu.Context(pr_precision=p, termcheck_threshold=1e-9 * g.n) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        for b in range(2):
            st = ctx.pr_run()
            r, d = ctx.pr_download()
            host = signed_mass_defect(off, r, d, DAMPING)
            assert abs(st["mass_defect"] - host) < 1e-13, (st["mass_defect"], host)
            assert abs(st["mass_total"] - (r.sum() + d.sum())) < 1e-13
            assert abs(st["mass_defect"]) < (1e-13 if prec == "fp64" else 1e-6)
            ops = g.gen_batch(seed=5, batch_id=b, n_del=200, n_ins=200)
            ctx.apply_batch(ops)
            g, _ = g.apply_batch(ops)
            off, col, _ = g.csr()


def test_nan_in_state_reports_numeric(gpu):
    """A NaN planted in the delta vector comes back as SUMINC_E_NUMERIC, not a hang
    or a silent result."""
    g = oracle.Graph.rmat(10, 16, seed=1)
    off, col, _ = g.csr()
    r0 = np.zeros(g.n)
    d0 = np.full(g.n, (1 - DAMPING) / g.n)
    d0[17] = np.nan
    for prec in (gpu.PR_FP64, gpu.PR_FP32_KAHAN):
        with gpu.Context(pr_precision=prec, termcheck_threshold=1e-9) as ctx:
            ctx.upload_csr(off, col)
            ctx.pr_upload(r0, d0)
            with pytest.raises(gpu.SumincError) as e:
                ctx.pr_run()
            assert e.value.code == gpu.E_NUMERIC
            d0[17] = np.inf
            ctx.pr_upload(r0, d0)
            with pytest.raises(gpu.SumincError) as e:
                ctx.pr_run()
            assert e.value.code == gpu.E_NUMERIC
            d0[17] = np.nan


def test_oversized_allocation_reports_nomem(gpu):
    """A device allocation beyond the GPU's memory comes back as SUMINC_E_NOMEM and
    leaves the context usable."""
    import ctypes as C

    g = oracle.Graph.rmat(10, 16, seed=1)
    off, col, _ = g.csr()
    with gpu.Context(pr_precision=gpu.PR_FP64, termcheck_threshold=1e-9) as ctx:
        ptr = C.c_void_p()
        rc = ctx._lib.suminc_device_alloc(ctx._h, 1 << 50, C.byref(ptr))
        assert rc == gpu.E_NOMEM and not ptr.value
        ctx.upload_csr(off, col)
        ctx.pr_init()
        assert ctx.pr_run()["residual"] < 1e-9


@pytest.mark.parametrize("layout", ["tiled", "binned"])
def test_tiled_pull_follows_precision_and_growth(gpu, layout):
    """The tiled and binned layouts are rebuilt when the precision changes (their
    source windows are sized in bytes of contributions) and after batches that grow
    the graph (the layout covers the new CSC, not the old edge count)."""
    g = oracle.Graph.rmat(14, 16, seed=2)
    off, col, _ = g.csr()
    T = 1e-10 * g.n
    kw = {"pull_tile_kib": 1} if layout == "tiled" else {"pull_bin_kib": 1}
    with gpu.Context(pr_precision=gpu.PR_FP32_KAHAN, termcheck_threshold=T, **kw) as ctx:
        ctx.upload_csr(off, col)
        for prec in (gpu.PR_FP32_KAHAN, gpu.PR_FP64, gpu.PR_FP32_KAHAN):
            ctx.set_params(pr_precision=prec)
            ctx.pr_init()
            st = ctx.pr_run()
            r, _ = ctx.pr_download()
            rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
            tol = 1e-11 if prec == gpu.PR_FP64 else 2e-6
            assert np.abs(r - rs).sum() <= st["residual"] / (1 - DAMPING) + tol, prec
        for b in range(3):  # three times as many inserts as deletes: m grows
            ops = g.gen_batch(seed=11, batch_id=b, n_del=int(0.005 * g.m), n_ins=int(0.015 * g.m))
            ctx.apply_batch(ops)
            g, _ = g.apply_batch(ops)
            st = ctx.pr_run()
            r, _ = ctx.pr_download()
            rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
            assert np.abs(r - rs).sum() <= st["residual"] / (1 - DAMPING) + 2e-6, b
        assert ctx.graph_info()[1] > len(col)


@pytest.mark.parametrize("path", ["plain", "sketch", "tiled", "binned", "sketch-binned"])
def test_incremental_fp32_meets_l1_target(gpu, path):
    """The accuracy bar after every incremental FP32/Kahan batch: L1 <= 1e-6 against
    the FP64 fixed point of the updated graph, at the benchmark's stop rule T = 1e-7,
    on the plain path, the SumInc query over the sketch, the tiled and the binned
    pull."""
    g = oracle.Graph.rmat(14, 16, seed=6)
    off, col, _ = g.csr()
    kw = {"pull_tile_kib": 1} if path == "tiled" else {"pull_bin_kib": 1} if path.endswith("binned") else {}
    with gpu.Context(pr_precision=gpu.PR_FP32_KAHAN, termcheck_threshold=1e-7, **kw) as ctx:
        ctx.upload_csr(off, col)
        if path.startswith("sketch"):
            ctx.cluster_leaves(1024, 2)
            ctx.sum_build(2, 1024, 1.0)
        run = ctx.sum_pr_run if path.startswith("sketch") else ctx.pr_run
        ctx.pr_init()
        run()
        for b in range(3):
            n_ops = int(0.01 * g.m)
            ops = g.gen_batch(seed=6, batch_id=b, n_del=n_ops // 2, n_ins=n_ops - n_ops // 2)
            ctx.apply_batch(ops)
            g, _ = g.apply_batch(ops)
            st = run()
            r, _ = ctx.pr_download()
            rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
            err = float(np.abs(r - rs).sum())
            assert err <= 1e-6, (path, b, err, st["residual"])
            assert abs(st["mass_defect"]) < 1e-6
