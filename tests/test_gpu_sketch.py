"""Supernode sketch on the device vs the oracle: clustering, acceptance and node
types bit-exact; shortcut rows vs the dense closed form; the SumInc PageRank
query (precompute, sketch rounds, correction, residual phase) vs the FP64 fixed
point, from scratch and after update batches."""
import numpy as np
import pytest

import oracle
from tests.conftest import DAMPING
from tests.test_oracle_sketch import shortcut_dense

pytestmark = pytest.mark.gpu
INF = 0xFFFFFFFF


def blocks(n, upto, size):
    return np.where(np.arange(n) < upto, np.arange(n) // size, INF).astype(np.uint32)


@pytest.mark.parametrize("scale,max_node,min_node", [(12, 1024, 2), (14, 4, 3), (16, 1024, 2)])
def test_leaf_clusters_equal_oracle(gpu, scale, max_node, min_node):
    g = oracle.Graph.rmat(scale, 16, seed=2)
    off, col, _ = g.csr()
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col)
        raw = ctx.cluster_leaves(max_node, min_node)
    assert np.array_equal(raw, g.leaf_clusters(max_node, min_node))


@pytest.mark.parametrize("kind,thr", [("leaf", 1.0), ("blocks", 1.0), ("blocks", -1e9)])
def test_structure_equals_oracle(gpu, kind, thr):
    g = oracle.Graph.rmat(13, 16, seed=1)
    off, col, _ = g.csr()
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col)
        if kind == "leaf":
            raw = ctx.cluster_leaves(1024, 2)
        else:
            raw = blocks(g.n, 4096, 8)
            ctx.set_clusters(raw)
        info = ctx.sum_build(2, 1024, thr)
        cid, typ = ctx.sum_download()
    ocid, otyp = g.sum_structure(raw, 2, 1024, thr)
    assert np.array_equal(cid, ocid)
    assert np.array_equal(typ, otyp)
    assert info["node_type_count"] == np.bincount(otyp, minlength=5).tolist()
    assert info["clusters"] == len(np.unique(ocid[ocid != INF]))
    assert info["entries"] == int(np.count_nonzero((otyp == 2) | (otyp == 3)))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_shortcut_rows_equal_dense_solve(gpu, prec):
    g = oracle.Graph.rmat(12, 16, seed=1)
    off, col, _ = g.csr()
    p = gpu.PR_FP64 if prec == "fp64" else gpu.PR_FP32_KAHAN
    with gpu.Context(pr_precision=p, termcheck_threshold=1e-30, shortcut_max_rounds=3000) as ctx:
        ctx.upload_csr(off, col)
        ctx.set_clusters(blocks(g.n, 1024, 16))
        ctx.sum_build(2, 1024, -1e9)
        cid, typ = ctx.sum_download()
        ctx.pr_init()
        entries = np.nonzero((typ == 2) | (typ == 3))[0]
        assert len(entries) > 20
        for s in entries[:20]:
            mem, reach, resid, nexit = ctx.sum_pr_shortcut(int(s))
            omem, x = shortcut_dense(g, cid, int(s), DAMPING)
            order = np.argsort(mem)
            assert np.array_equal(mem[order], omem)
            tol = 1e-12 if prec == "fp64" else 2e-6
            assert np.abs(reach[order] - x).max() < tol
            assert np.abs(resid).sum() < tol
            # exits come first in the row
            assert np.all(np.isin(typ[mem[:nexit]], [1, 3])) and not np.any(np.isin(typ[mem[nexit:]], [1, 3]))


@pytest.mark.parametrize("compact", [0, 1])
@pytest.mark.parametrize("kind", ["leaf", "blocks"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_sum_pr_query_equals_fixed_point(gpu, kind, prec, compact):
    g = oracle.Graph.rmat(14, 16, seed=1)
    off, col, _ = g.csr()
    rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
    T = 1e-10 * g.n
    p = gpu.PR_FP64 if prec == "fp64" else gpu.PR_FP32_KAHAN
    with gpu.Context(pr_precision=p, termcheck_threshold=T, pull_compact_kib=compact) as ctx:
        ctx.upload_csr(off, col)
        if kind == "leaf":
            ctx.cluster_leaves(1024, 2)
            ctx.sum_build(2, 1024, 1.0)
        else:
            ctx.set_clusters(blocks(g.n, 8192, 8))
            ctx.sum_build(2, 1024, -1e9)
        ctx.pr_init()
        st = ctx.sum_pr_run()
        r, d = ctx.pr_download()
    err = np.abs(r - rs).sum()
    assert err <= st["residual"] / (1 - DAMPING) + (1e-11 if prec == "fp64" else 2e-6), (err, st)
    assert st["sketch_rounds"] > 0  # sketch rounds ran


# compact=1: both pulls gather from compact source positions (the default only on
# large graphs), the residual phase in 16 KiB source bands over positions in id
# order; compact=2: positions by decreasing out-degree class, unbanded (the large-graph
# default), so the sketch rounds' skip test reads the cluster by position;
# compact=3: the same positions under the tiled pull (1 KiB source windows);
# compact=4: under the binned pull (1 KiB windows)
@pytest.mark.parametrize("prec,compact", [("fp64", 0), ("fp32", 0), ("fp32", 1), ("fp32", 2), ("fp64", 2), ("fp32", 3),
                                          ("fp64", 3), ("fp32", 4), ("fp64", 4)])
def test_sum_pr_incremental(gpu, prec, compact):
    g = oracle.Graph.rmat(13, 16, seed=3)
    off, col, _ = g.csr()
    T = 1e-10 * g.n
    p = gpu.PR_FP64 if prec == "fp64" else gpu.PR_FP32_KAHAN
    kw = {1: {"pull_compact_kib": 1, "pull_band_kib": 16}, 2: {"pull_compact_kib": 1},
          3: {"pull_compact_kib": 1, "pull_tile_kib": 1}, 4: {"pull_compact_kib": 1, "pull_bin_kib": 1}}.get(compact, {})
    with gpu.Context(pr_precision=p, termcheck_threshold=T, **kw) as ctx:
        ctx.upload_csr(off, col)
        ctx.set_clusters(blocks(g.n, 4096, 8))
        ctx.sum_build(2, 1024, -1e9)
        ctx.pr_init()
        ctx.sum_pr_run()
        for b in range(3):
            n_ops = int(0.01 * g.m)
            ops = g.gen_batch(seed=3, batch_id=b, n_del=n_ops // 2, n_ins=n_ops - n_ops // 2)
            ctx.apply_batch(ops)
            g, _ = g.apply_batch(ops)
            st = ctx.sum_pr_run()
            r, _ = ctx.pr_download()
            rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
            err = np.abs(r - rs).sum()
            assert err <= st["residual"] / (1 - DAMPING) + (1e-11 if prec == "fp64" else 2e-6), (b, err, st)
        # the structure follows the new graph, acceptance is kept
        cid, typ = ctx.sum_download()
    ocid, otyp = g.sum_structure(blocks(g.n, 4096, 8), 2, 1024, -1e9)
    assert np.array_equal(cid, ocid) and np.array_equal(typ, otyp)


def test_sum_build_errors(gpu):
    g = oracle.Graph.rmat(10, 8, seed=1)
    off, col, _ = g.csr()
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col)
        with pytest.raises(gpu.SumincError) as e:
            ctx.sum_build()
        assert e.value.code == -4  # no cluster assignment yet
        bad = np.zeros(g.n, np.uint32)
        bad[0] = g.n + 5
        with pytest.raises(gpu.SumincError):
            ctx.set_clusters(bad)
        with pytest.raises(gpu.SumincError):
            ctx.cluster_leaves(5000, 2)  # beyond the 2048-member LDS bound


def _sssp_graph(weights, seed=2, scale=13):
    return oracle.Graph.rmat(scale, 16, seed=seed, weights=weights, wmin=1, wmax=64)


def _upload(ctx, g):
    off, col, w = g.csr()
    ctx.upload_csr(off, col, w)


def _cluster(ctx, g, kind):
    if kind == "leaf":
        ctx.cluster_leaves(1024, 2)
        return ctx.sum_build(2, 1024, 1.0)
    ctx.set_clusters(blocks(g.n, 4096, 8))
    return ctx.sum_build(2, 1024, -1e9)


@pytest.mark.parametrize("weights", [None, "u32"])
def test_sssp_shortcut_rows_equal_oracle(gpu, weights):
    g = _sssp_graph(weights, scale=12)
    off, col, w = g.csr()
    wv = np.ones(len(col), np.uint64) if w is None else w.astype(np.uint64)
    with gpu.Context() as ctx:
        _upload(ctx, g)
        _cluster(ctx, g, "blocks")
        cid, typ = ctx.sum_download()
        entries = np.nonzero((typ == 2) | (typ == 3))[0]
        assert len(entries) > 20
        for s in entries[:40]:
            mem, dist, par, nexit = ctx.sum_sssp_shortcut(int(s))
            exp = g.sum_sssp_shortcut(cid, int(s))
            # rows keep the exits (the first nexit members)
            assert np.array_equal(dist[:nexit], exp[mem[:nexit]]), int(s)
            assert np.all(dist[nexit:] == INF) and np.all(par[nexit:] == INF)
            # parent: smallest in-cluster predecessor on a shortest in-cluster path
            for j, x in enumerate(mem[:nexit]):
                if x == s or exp[x] == INF:
                    assert par[j] == INF
                    continue
                cands = [u for u in mem for e in range(off[u], off[u + 1])
                         if col[e] == x and exp[u] != INF and int(exp[u]) + int(wv[e]) == int(exp[x])]
                assert par[j] == min(cands), (int(s), int(x))
            assert np.all(np.isin(typ[mem[:nexit]], [1, 3]))


@pytest.mark.parametrize("weights", [None, "u32"])
@pytest.mark.parametrize("kind", ["leaf", "blocks"])
def test_sum_sssp_equals_dijkstra(gpu, weights, kind):
    g = _sssp_graph(weights)
    with gpu.Context() as ctx:
        _upload(ctx, g)
        _cluster(ctx, g, kind)
        cid, typ = ctx.sum_download()
        sources = [int(np.nonzero(typ == 0)[0][0])]
        for t in (1, 2, 3, 4):
            if np.any(typ == t):
                sources.append(int(np.nonzero(typ == t)[0][0]))
        for s in sources:
            ctx.sssp_init(s)
            st = ctx.sum_sssp_run()
            dist, par = ctx.sssp_download()
            od, op = g.dijkstra(s)
            assert np.array_equal(dist, od), (s, int(typ[s]))
            assert np.array_equal(par, op), (s, int(typ[s]))
            assert st["rounds"] >= 1


@pytest.mark.parametrize("weights", [None, "u32"])
@pytest.mark.parametrize("kind", ["leaf", "blocks"])
def test_sum_sssp_incremental(gpu, weights, kind):
    g = _sssp_graph(weights, seed=5)
    src = 0
    with gpu.Context() as ctx:
        _upload(ctx, g)
        _cluster(ctx, g, kind)
        ctx.sssp_init(src)
        ctx.sum_sssp_run()
        for b in range(4):
            n_ops = int(0.01 * g.m)
            ops = g.gen_batch(seed=5, batch_id=b, n_del=n_ops // 2, n_ins=n_ops - n_ops // 2)
            ctx.apply_batch(ops)
            g, _ = g.apply_batch(ops)
            ctx.sum_sssp_run()
            dist, par = ctx.sssp_download()
            od, op = g.dijkstra(src)
            assert np.array_equal(dist, od), b
            assert np.array_equal(par, op), b


def test_sum_sssp_mixed_with_plain_runs(gpu):
    """Plain and sketch runs share one state: alternating them across batches keeps
    the exact result."""
    g = _sssp_graph("u32", seed=7)
    with gpu.Context() as ctx:
        _upload(ctx, g)
        _cluster(ctx, g, "blocks")
        ctx.sssp_init(3)
        ctx.sssp_run()
        for b in range(4):
            ops = g.gen_batch(seed=7, batch_id=b, n_del=300, n_ins=300)
            ctx.apply_batch(ops)
            g, _ = g.apply_batch(ops)
            (ctx.sum_sssp_run if b % 2 == 0 else ctx.sssp_run)()
            dist, par = ctx.sssp_download()
            od, op = g.dijkstra(3)
            assert np.array_equal(dist, od) and np.array_equal(par, op), b


def test_zero_integer_weight_rejected(gpu):
    with gpu.Context() as ctx:
        with pytest.raises(gpu.SumincError):
            ctx.upload_csr(np.array([0, 1, 1], np.uint64), np.array([1], np.uint32), np.array([0], np.uint32))
        ctx.upload_csr(np.array([0, 1, 1], np.uint64), np.array([1], np.uint32), np.array([3], np.uint32))
        ops = np.zeros(1, gpu.UPDATE_DTYPE)
        ops[0] = (gpu.OP_ADD, 1, 0, 0, 0.0, 0)
        with pytest.raises(gpu.SumincError):
            ctx.apply_batch(ops)


def test_rows_follow_small_batches(gpu):
    """Small batches leave most clusters untouched: their shortcut rows are moved
    to the new layout, the touched ones are recomputed.  Both kinds must equal the
    oracle's rows on the new graph, and both queries stay exact."""
    g = _sssp_graph("u32", seed=8, scale=12)
    T = 1e-10 * g.n
    with gpu.Context(pr_precision=gpu.PR_FP64, termcheck_threshold=T) as ctx:
        _upload(ctx, g)
        ctx.set_clusters(blocks(g.n, 4096, 8))
        ctx.sum_build(2, 1024, -1e9)
        ctx.sssp_init(0)
        ctx.sum_sssp_run()
        ctx.pr_init()
        ctx.sum_pr_run()
        for b in range(3):
            ops = g.gen_batch(seed=8, batch_id=b, n_del=10, n_ins=10)
            ctx.apply_batch(ops)
            g, _ = g.apply_batch(ops)
            ctx.sum_sssp_run()
            dist, par = ctx.sssp_download()
            od, op = g.dijkstra(0)
            assert np.array_equal(dist, od) and np.array_equal(par, op), b
            st = ctx.sum_pr_run()
            r, _ = ctx.pr_download()
            rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
            assert np.abs(r - rs).sum() <= st["residual"] / (1 - DAMPING) + 1e-11, b
            cid, typ = ctx.sum_download()
            entries = np.nonzero((typ == 2) | (typ == 3))[0]
            for s in entries[::max(1, len(entries) // 40)]:
                mem, sd, _, ne = ctx.sum_sssp_shortcut(int(s))
                assert np.array_equal(sd[:ne], g.sum_sssp_shortcut(cid, int(s))[mem[:ne]]), (b, int(s))
                mem, reach, _, _ = ctx.sum_pr_shortcut(int(s))
                omem, x = shortcut_dense(g, cid, int(s), DAMPING)
                order = np.argsort(mem)
                assert np.abs(reach[order] - x).max() < 1e-6, (b, int(s))


def _canonical(cid):
    """Cluster labels replaced by the smallest member id (INF stays): partitions
    compare equal whatever their numbering."""
    out = np.full(len(cid), INF, np.uint64)
    m = cid != INF
    if m.any():
        ids = np.nonzero(m)[0]
        first = {}
        for v, k in zip(ids, cid[m]):
            first.setdefault(int(k), int(v))
        out[m] = [first[int(k)] for k in cid[m]]
    return out


def test_degraded_clusters_dissolve(gpu):
    """Batches that give a cluster so many entries and exits that its benefit falls
    to the build's threshold dissolve it (its members become Singles and stay so).
    The expected structure after every batch: the oracle's acceptance on the
    current graph over the clusters that survived every earlier batch.  The SumInc
    PageRank and SSSP queries stay equal to the fixed point and to Dijkstra."""
    g = oracle.Graph.rmat(13, 16, seed=6, weights="u32", wmin=1, wmax=64)
    off, col, w = g.csr()
    raw = g.leaf_clusters(1024, 2)
    with gpu.Context(pr_precision=gpu.PR_FP64, termcheck_threshold=1e-10 * g.n) as ctx:
        ctx.upload_csr(off, col, w)
        ctx.set_clusters(raw)
        info0 = ctx.sum_build(2, 1024, 1.0)
        assert info0["dissolved_clusters"] == 0
        ctx.pr_init()
        ctx.sum_pr_run()
        ctx.sssp_init(0)
        ctx.sum_sssp_run()
        ocid0, _ = g.sum_structure(raw, 2, 1024, 1.0)
        alive = np.where(ocid0 != INF, raw, INF).astype(np.uint32)  # accepted by the build
        for b in range(4):
            n_ops = int(0.05 * g.m)
            ops = g.gen_batch(seed=6, batch_id=b, n_del=n_ops // 2, n_ins=n_ops - n_ops // 2, wmin=1, wmax=64)
            ctx.apply_batch(ops)
            g, _ = g.apply_batch(ops)
            st = ctx.sum_pr_run()
            ctx.sum_sssp_run()
            cid, typ = ctx.sum_download()
            info = ctx.sum_info()
            ocid, otyp = g.sum_structure(alive, 2, 1024, 1.0)
            assert np.array_equal(_canonical(cid), _canonical(ocid)), b
            assert np.array_equal(typ, otyp), b
            alive = np.where(ocid != INF, raw, INF).astype(np.uint32)  # the survivors, by raw id
            r, _ = ctx.pr_download()
            rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
            assert np.abs(r - rs).sum() <= st["residual"] / (1 - DAMPING) + 1e-11
            dist, par = ctx.sssp_download()
            od, op = g.dijkstra(0)
            assert np.array_equal(dist, od) and np.array_equal(par, op)
        assert info["dissolved_clusters"] > 0
        assert info["clusters"] + info["dissolved_clusters"] == info0["clusters"]
