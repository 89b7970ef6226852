"""The HIP path on every workload of the metric, at its full size.

C1 (R-MAT s16 seed 1, plain FP64 delta-PageRank, per-vertex threshold 1e-9) runs
at its exact size in test_gpu_pr.py::test_pagerank_vs_fp64_oracle.  C2 .. C5
below run at their own scale, seed, batch fraction and precision:

  C2  s22 seed 2, u32 weights U[1,64], leaf-cluster supernodes (<= 1024 members),
      three 0.1% batches: the device graph equals the oracle's graph after the same
      batches, and the SumInc SSSP query equals the oracle's Dijkstra bit for bit
      (distances and parents).
  C3  s24 seed 3, one 1% batch, FP32 deltas + Kahan ranks + FP64 residual, over
      the plain CSR and over the supernode sketch: L1 <= 1e-6 against the FP64 run
      on the same graph (the metric's tolerance), and the FP64 mass balance.
  C4  s22 seed 4, FP32 weights U(0,1], Dirichlet(1) personalization over 1024
      seeds, loss sum c_i r_i with c ~ N(0,1): the FP32 adjoint gradient against
      central differences (h = 1e-3) of the FP64 forward on 256 random edges
      (1e-4 relative, floored at 1% of the largest difference) and against the
      FP64 adjoint.
  C5  s26 seed 5, ten successive 1% batches (SSSP weights U[1,255]): after every
      batch the SumInc PageRank query's mass balance; at the end SSSP equals a
      plain full recompute bit for bit and PageRank is within L1 <= 1e-6 of the
      FP64 run.  The adjoint on the same graph with FP32 weights: FP32 against
      FP64 gradients and central differences on random edges.

The FP64 reference of the large cases is the device's own FP64 delta-PageRank run
to sum|delta| < 1e-12 or below; that path is pinned against scipy's sparse solve
of (I - d P^T) r = (1 - d)/N in test_gpu_pr.py.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import DAMPING

pytestmark = pytest.mark.gpu


def _batch_ops(ctx, seed, b, frac, wmin=1, wmax=64):
    m = ctx.graph_info()[1]
    n_ops = int(frac * m)
    return ctx.generate_batch(seed, b, n_ops // 2, n_ops - n_ops // 2, wmin, wmax)


@pytest.mark.timeout(400)
def test_c2_sssp_s22_supernodes_bit_exact(gpu):
    g = oracle.Graph.rmat(22, 16, seed=2, weights="u32", wmin=1, wmax=64)
    with gpu.Context(max_rounds=1_000_000) as ctx:
        ctx.generate_rmat(22, 16, 0.57, 0.19, 0.19, seed=2, weights="u32", wmin=1, wmax=64)
        n, m, _ = ctx.graph_info()
        assert (n, m) == (g.n, g.m)
        ctx.cluster_leaves(1024, 2)
        info = ctx.sum_build(2, 1024, 1.0)
        assert info["clusters"] > 0 and info["max_cluster"] <= 1024
        ctx.sssp_init(0)
        ctx.sum_sssp_run()
        for b in range(3):
            ops = _batch_ops(ctx, 2, b, 0.001)
            dev = ctx.upload_batch(ops)
            ctx.apply_batch_device(dev)
            dev.free()
            st = ctx.sum_sssp_run()
            assert st["rounds"] >= 1
            g, _ = g.apply_batch(ops)
        dist, par = ctx.sssp_download()
        off, col, w = ctx.download_csr()
    go, gc, gw = g.csr()
    assert np.array_equal(off, go) and np.array_equal(col, gc) and np.array_equal(w, gw)
    od, op = g.dijkstra(0)
    assert np.array_equal(dist, od), int((dist != od).sum())
    assert np.array_equal(par, op), int((par != op).sum())


@pytest.mark.timeout(300)
@pytest.mark.parametrize("sketch", [False, True], ids=["plain", "supernodes"])
def test_c3_pagerank_s24_fp32_kahan_l1(gpu, sketch):
    T = 1e-7  # sum|delta| < T bounds the truncation part of the L1 error by T / (1 - d)
    with gpu.Context(termcheck_threshold=T, pr_precision=gpu.PR_FP32_KAHAN, max_rounds=1_000_000) as ctx:
        ctx.generate_rmat(24, 16, 0.57, 0.19, 0.19, seed=3)
        if sketch:
            ctx.cluster_leaves(1024, 2)
            assert ctx.sum_build(2, 1024, 1.0)["clusters"] > 0
        run = ctx.sum_pr_run if sketch else ctx.pr_run
        ctx.pr_init()
        run()
        ctx.apply_batch(_batch_ops(ctx, 3, 0, 0.01))
        st = run()
        r32, _ = ctx.pr_download()
        assert st["residual"] < T
        assert abs(st["mass_defect"]) < 1e-5, st
        ctx.set_params(pr_precision=gpu.PR_FP64, termcheck_threshold=1e-13)
        ctx.pr_init()
        st64 = ctx.pr_run()
        r64, d64 = ctx.pr_download()
        assert abs(st64["mass_defect"]) < 1e-12, st64
    l1 = float(np.abs(r32 - r64).sum())
    linf = float(np.abs(r32 - r64).max())
    print(f"C3 {'sketch' if sketch else 'plain'}: L1 {l1:.3e} Linf {linf:.3e} mass defect fp32 "
          f"{st['mass_defect']:.3e} fp64 {st64['mass_defect']:.3e}")
    assert l1 <= 1e-6, (l1, linf)


@pytest.mark.timeout(400)
def test_c4_adjoint_s22_fd_256_edges(gpu):
    rng = np.random.default_rng(4)
    h = 1e-3
    with gpu.Context(termcheck_threshold=1e-9, pr_precision=gpu.PR_FP32_KAHAN, max_rounds=1_000_000) as ctx:
        ctx.generate_rmat(22, 16, 0.57, 0.19, 0.19, seed=4, weights="f32")
        n, m, _ = ctx.graph_info()
        p = np.zeros(n, np.float32)
        seeds = rng.choice(n, 1024, replace=False)
        p[seeds] = rng.dirichlet(np.ones(1024)).astype(np.float32)
        c = rng.standard_normal(n, dtype=np.float32)
        off, col, w = ctx.download_csr()
        ctx.wpr_forward(p)
        gp32, gw32, _ = ctx.wpr_backward(c)
    with gpu.Context(termcheck_threshold=1e-15, pr_precision=gpu.PR_FP64, max_rounds=1_000_000) as c64:
        c64.upload_csr(off, col, w)
        c64.wpr_forward(p)
        gp64, gw64, _ = c64.wpr_backward(c)
        assert np.abs(gw32 - gw64).max() <= 1e-4 * np.abs(gw64).max()
        assert np.abs(gp32 - gp64).max() <= 1e-4 * np.abs(gp64).max()
        # central differences of the FP64 forward, the weights kept in device memory
        # and one element rewritten per evaluation
        wd = c64.device_array(w)
        cd = c.astype(np.float64)
        edges = rng.choice(m, 256, replace=False)
        fd = np.zeros(len(edges))
        for i, e in enumerate(edges):
            w0 = w[e]
            wp, wm = np.float32(w0 + np.float32(h)), np.float32(w0 - np.float32(h))
            wd.set(e, wp)
            c64.wpr_set_weights(wd)
            c64.wpr_forward(p)
            lp = cd @ c64.wpr_download()[0]
            wd.set(e, wm)
            c64.wpr_set_weights(wd)
            c64.wpr_forward(p)
            lm = cd @ c64.wpr_download()[0]
            wd.set(e, w0)
            fd[i] = (lp - lm) / (np.float64(wp) - np.float64(wm))
        wd.free()
    ga = gw32[edges].astype(np.float64)
    rel = np.abs(fd - ga) / np.maximum(np.abs(fd), 1e-2 * np.abs(fd).max())
    print(f"C4: FD on {len(edges)} edges, max rel err {rel.max():.3e}; |fd| max {np.abs(fd).max():.3e}")
    assert np.abs(fd).max() > 0
    assert rel.max() <= 1e-4, rel.max()


@pytest.mark.timeout(400)
def test_c5_web26_ten_batches(gpu):
    T = 1e-7
    with gpu.Context(termcheck_threshold=T, pr_precision=gpu.PR_FP32_KAHAN, max_rounds=1_000_000) as ctx:
        ctx.generate_rmat(26, 16, 0.57, 0.19, 0.19, seed=5, weights="u32", wmin=1, wmax=255)
        ctx.cluster_leaves(1024, 2)
        assert ctx.sum_build(2, 1024, 1.0)["clusters"] > 0
        ctx.pr_init()
        ctx.sum_pr_run()
        ctx.sssp_init(0)
        ctx.sum_sssp_run()
        defects = []
        for b in range(10):
            dev = ctx.upload_batch(_batch_ops(ctx, 5, b, 0.01, 1, 255))
            ctx.apply_batch_device(dev)
            dev.free()
            st = ctx.sum_pr_run()
            assert st["residual"] < T
            defects.append(st["mass_defect"])
            ctx.sum_sssp_run()
        print(f"C5 mass defects per batch: {['%.2e' % x for x in defects]}")
        assert max(abs(x) for x in defects) < 1e-5
        r32, _ = ctx.pr_download()
        dist, par = ctx.sssp_download()
        ctx.sssp_init(0)
        ctx.sssp_run()
        d2, p2 = ctx.sssp_download()
        assert np.array_equal(dist, d2), int((dist != d2).sum())
        assert np.array_equal(par, p2), int((par != p2).sum())
        del d2, p2, dist, par
        ctx.set_params(pr_precision=gpu.PR_FP64, termcheck_threshold=1e-12)
        ctx.pr_init()
        ctx.pr_run()
        r64, _ = ctx.pr_download()
    l1 = float(np.abs(r32 - r64).sum())
    print(f"C5: L1 {l1:.3e} Linf {float(np.abs(r32 - r64).max()):.3e}")
    assert l1 <= 1e-6


@pytest.mark.timeout(400)
def test_c5_adjoint_s26(gpu):
    rng = np.random.default_rng(5)
    h = 1e-3
    with gpu.Context(termcheck_threshold=1e-9, pr_precision=gpu.PR_FP32_KAHAN, max_rounds=1_000_000) as ctx:
        ctx.generate_rmat(26, 16, 0.57, 0.19, 0.19, seed=5, weights="f32")
        n, m, _ = ctx.graph_info()
        p = np.zeros(n, np.float32)
        seeds = rng.choice(n, 1024, replace=False)
        p[seeds] = rng.dirichlet(np.ones(1024)).astype(np.float32)
        c = rng.standard_normal(n, dtype=np.float32)
        ctx.wpr_forward(p)
        _, gw32, _ = ctx.wpr_backward(c)
        ctx.set_params(pr_precision=gpu.PR_FP64, termcheck_threshold=1e-15)
        ctx.wpr_forward(p)
        _, gw64, _ = ctx.wpr_backward(c)
        assert np.abs(gw32 - gw64).max() <= 1e-4 * np.abs(gw64).max()
        # central differences on the edges with the largest gradients (a random
        # edge of the 1.07B mostly carries none at this personalization)
        off, col, w = ctx.download_csr()
        wd = ctx.device_array(w)
        cd = c.astype(np.float64)
        edges = np.argsort(-np.abs(gw64))[:4]
        fd = np.zeros(len(edges))
        for i, e in enumerate(edges):
            w0 = w[e]
            wp, wm = np.float32(w0 + np.float32(h)), np.float32(w0 - np.float32(h))
            wd.set(e, wp)
            ctx.wpr_set_weights(wd)
            ctx.wpr_forward(p)
            lp = cd @ ctx.wpr_download()[0]
            wd.set(e, wm)
            ctx.wpr_set_weights(wd)
            ctx.wpr_forward(p)
            lm = cd @ ctx.wpr_download()[0]
            wd.set(e, w0)
            fd[i] = (lp - lm) / (np.float64(wp) - np.float64(wm))
        wd.free()
    rel = np.abs(fd - gw32[edges]) / np.abs(fd)
    print(f"C5 adjoint: FD on {len(edges)} edges, rel err {rel}")
    assert rel.max() <= 1e-4
