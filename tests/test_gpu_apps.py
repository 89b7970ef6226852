"""BFS, SSWP and PHP on the device against the oracle: BFS hops and SSWP widths
bit-exact (from scratch, one step at a time, and after update batches through
the batch repair), parents valid and acyclic; PHP against the FP64 closed form."""
import numpy as np
import pytest

import oracle
from oracle import apps

pytestmark = pytest.mark.gpu
INF = 0xFFFFFFFF
D_PHP = float(np.float32(0.8))


def _graph(weights, scale=13, seed=2):
    return oracle.Graph.rmat(scale, 16, seed=seed, weights=weights, wmin=1, wmax=255)


def _check_parents(app, off, col, w, src, val, parent):
    """Every reached vertex's parent has an edge to it that yields its value, and
    following parents reaches the source without a cycle."""
    n = len(off) - 1
    for v in range(n):
        if v == src or (val[v] == (0 if app == "sswp" else INF)):
            continue
        p = int(parent[v])
        row = range(int(off[p]), int(off[p + 1]))
        ok = False
        for e in row:
            if col[e] != v:
                continue
            if app == "bfs":
                ok |= int(val[p]) + 1 == int(val[v])
            else:
                ok |= min(int(val[p]), int(w[e])) == int(val[v])
        assert ok, (v, p)
    # acyclic: walk up from every vertex at most n steps
    depth = np.full(n, -1)
    depth[src] = 0
    for v in range(n):
        if depth[v] >= 0 or val[v] == (0 if app == "sswp" else INF):
            continue
        chain = []
        x = v
        while depth[x] < 0:
            chain.append(x)
            x = int(parent[x])
            assert len(chain) <= n, "parent cycle"
        for k, y in enumerate(reversed(chain)):
            depth[y] = depth[x] + k + 1


def _expected(app, off, col, w, src):
    return apps.bfs(off, col, src) if app == "bfs" else apps.sswp(off, col, w, src)


@pytest.mark.parametrize("app", ["bfs", "sswp"])
def test_traversal_apps_equal_oracle(gpu, app):
    g = _graph("u32")
    off, col, w = g.csr()
    code = gpu.APP_BFS if app == "bfs" else gpu.APP_SSWP
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col, w)
        ctx.traversal_init(code, 5)
        st = ctx.sssp_run()
        val, par = ctx.sssp_download()
        # one step at a time gives the same values
        ctx.traversal_init(code, 5)
        steps = 0
        while ctx.sssp_step():
            steps += 1
        val2, _ = ctx.sssp_download()
    exp = _expected(app, off, col, w, 5)
    assert np.array_equal(val, exp)
    assert np.array_equal(val2, exp)
    assert st["rounds"] > 2 and steps > 2
    _check_parents(app, off, col, w, 5, val, par)
    if app == "sswp":
        assert val[5] == INF and np.count_nonzero(val) > g.n // 4


@pytest.mark.parametrize("app", ["bfs", "sswp"])
def test_traversal_apps_after_batches(gpu, app):
    g = _graph("u32", scale=12, seed=4)
    off, col, w = g.csr()
    code = gpu.APP_BFS if app == "bfs" else gpu.APP_SSWP
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col, w)
        ctx.traversal_init(code, 0)
        ctx.sssp_run()
        for b in range(4):
            n_ops = int(0.01 * g.m)
            ops = g.gen_batch(seed=9, batch_id=b, n_del=n_ops // 2, n_ins=n_ops - n_ops // 2, wmin=1, wmax=255)
            ctx.apply_batch(ops)
            g, _ = g.apply_batch(ops)
            ctx.sssp_run()
            val, par = ctx.sssp_download()
            off, col, w = g.csr()
            assert np.array_equal(val, _expected(app, off, col, w, 0)), b
            _check_parents(app, off, col, w, 0, val, par)


def test_bfs_ignores_weights_and_sketch_refuses_other_apps(gpu):
    g = _graph("u32", scale=11)
    off, col, w = g.csr()
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col, w)
        ctx.traversal_init(gpu.APP_BFS, 0)
        ctx.sssp_run()
        val, _ = ctx.sssp_download()
        assert np.array_equal(val, apps.bfs(off, col, 0))
        ctx.cluster_leaves(1024, 2)
        ctx.sum_build(2, 1024, 1.0)
        ctx.traversal_init(gpu.APP_SSWP, 0)
        with pytest.raises(gpu.SumincError) as e:
            ctx.sum_sssp_run()
        assert e.value.code == gpu.E_STATE
        with pytest.raises(gpu.SumincError) as e:
            ctx.traversal_init(7, 0)
        assert e.value.code == gpu.E_INVALID


# binned: the series rounds through the binned layout (1 KiB source windows)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("weights,binned", [(None, False), ("u32", False), ("f32", False), ("f32", True)])
def test_php_equals_closed_form(gpu, prec, weights, binned):
    g = oracle.Graph.rmat(13, 16, seed=7, weights=weights, wmin=1, wmax=64)
    off, col, w = g.csr()
    src = 3
    exp = apps.php(off, col, w, src, D_PHP)
    T = 1e-12 if prec == "fp64" else 1e-7
    p = gpu.PR_FP64 if prec == "fp64" else gpu.PR_FP32_KAHAN
    kw = {"pull_bin_kib": 1} if binned else {}
    with gpu.Context(pr_precision=p, termcheck_threshold=T, **kw) as ctx:
        ctx.upload_csr(off, col, w)
        val, st = ctx.php_run(src)
    assert val[src] == pytest.approx(1.0, abs=1e-6)
    err = np.abs(val - exp).sum()
    # the series stopped at sum |delta| < T: what is left adds at most T / (1 - d)
    assert err <= st["residual"] / (1 - D_PHP) + (1e-12 if prec == "fp64" else 2e-6 * exp.sum()), (err, st)
    assert st["rounds"] > 5
