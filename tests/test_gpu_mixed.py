"""PageRank and SSSP kept on one context across successive batches, queried in
either order: every SSSP result stays bit-exact vs Dijkstra and every PageRank
result within its residual bound of the FP64 fixed point.  The SSSP seeds a batch
leaves behind belong to the SSSP state, so a PageRank query (which resets the
context's round counters and reuses its scratch lists) between the batch and the
SSSP query must not lose them, and neither may a second batch applied before any
query."""
import numpy as np
import pytest

import oracle
from tests.conftest import DAMPING

pytestmark = pytest.mark.gpu


def check_pr(ctx, g, st):
    rank, _ = ctx.pr_download()
    rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
    err = float(np.abs(rank - rs).sum())
    assert err <= st["residual"] / (1 - DAMPING) + 2e-6


def check_sssp(ctx, g, src):
    d, p = ctx.sssp_download()
    od, op = g.dijkstra(src)
    assert np.array_equal(d, od)
    assert np.array_equal(p, op)


# compact=1: the plain pull over compact source positions in 16 KiB source bands
# (both used by default only on large graphs)
@pytest.mark.parametrize("sketch,compact", [(False, 0), (True, 0), (False, 1), (True, 1)])
def test_pagerank_between_batch_and_sssp(gpu, sketch, compact):
    g = oracle.Graph.rmat(14, 16, seed=5, weights="u32", wmin=1, wmax=255)
    off, col, w = g.csr()
    src = 0
    kw = {"pull_compact_kib": 1, "pull_band_kib": 16} if compact else {}
    with gpu.Context(termcheck_threshold=1e-9, **kw) as ctx:
        ctx.upload_csr(off, col, w)
        if sketch:
            ctx.cluster_leaves(1024, 2)
            assert ctx.sum_build(2, 1024, 1.0)["clusters"] > 0
        ss = ctx.sum_sssp_run if sketch else ctx.sssp_run
        ctx.sssp_init(src)
        ss()
        ctx.pr_init()
        ctx.pr_run()
        gcur = g
        n_ops = len(col) // 100
        for b in range(3):
            ops = gcur.gen_batch(seed=5, batch_id=b, n_del=n_ops // 2, n_ins=n_ops - n_ops // 2, wmin=1, wmax=255)
            gcur, _ = gcur.apply_batch(ops)
            ctx.apply_batch(ops)
            st = ctx.pr_run()  # PageRank first: the SSSP seeds of the batch must survive it
            st_ss = ss()
            assert st_ss["rounds"] > 0
            check_sssp(ctx, gcur, src)
            check_pr(ctx, gcur, st)


def test_two_batches_before_the_sssp_query(gpu):
    g = oracle.Graph.rmat(13, 16, seed=6, weights="u32", wmin=1, wmax=64)
    off, col, w = g.csr()
    with gpu.Context(termcheck_threshold=1e-9) as ctx:
        ctx.upload_csr(off, col, w)
        ctx.sssp_init(0)
        ctx.sssp_run()
        ctx.pr_init()
        ctx.pr_run()
        gcur = g
        n_ops = len(col) // 200
        for b in range(2):  # no query between the two batches
            ops = gcur.gen_batch(seed=6, batch_id=b, n_del=n_ops // 2, n_ins=n_ops - n_ops // 2, wmin=1, wmax=64)
            gcur, _ = gcur.apply_batch(ops)
            ctx.apply_batch(ops)
        st = ctx.pr_run()
        ctx.sssp_run()
        check_sssp(ctx, gcur, 0)
        check_pr(ctx, gcur, st)


def test_pagerank_after_sssp_only_batches(gpu):
    """Batches on an SSSP-only context leave the CSC stale; a PageRank state made
    afterwards rebuilds it, and its next batches merge it again."""
    g = oracle.Graph.rmat(13, 16, seed=7, weights="u32", wmin=1, wmax=64)
    off, col, w = g.csr()
    with gpu.Context(termcheck_threshold=1e-9) as ctx:
        ctx.upload_csr(off, col, w)
        ctx.sssp_init(0)
        ctx.sssp_run()
        gcur = g
        n_ops = len(col) // 100
        for b in range(2):
            ops = gcur.gen_batch(seed=7, batch_id=b, n_del=n_ops // 2, n_ins=n_ops - n_ops // 2, wmin=1, wmax=64)
            gcur, _ = gcur.apply_batch(ops)
            ctx.apply_batch(ops)
            ctx.sssp_run()
            check_sssp(ctx, gcur, 0)
        ctx.pr_init()
        st = ctx.pr_run()
        check_pr(ctx, gcur, st)
        for b in range(2, 4):
            ops = gcur.gen_batch(seed=7, batch_id=b, n_del=n_ops // 2, n_ins=n_ops - n_ops // 2, wmin=1, wmax=64)
            gcur, _ = gcur.apply_batch(ops)
            ctx.apply_batch(ops)
            st = ctx.pr_run()
            ctx.sssp_run()
            check_sssp(ctx, gcur, 0)
            check_pr(ctx, gcur, st)
