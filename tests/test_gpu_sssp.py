"""Min-plus SSSP on the device: distances and parents bit-exact vs Dijkstra
(parent = smallest-id predecessor on a shortest path), plain and incremental."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def run_sssp(gpu, off, col, w, src):
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col, w)
        ctx.sssp_init(src)
        st = ctx.sssp_run()
        d, p = ctx.sssp_download()
    return d, p, st


def test_golden_test_base(gpu):
    exp = json.loads((GOLD / "test_base_expected.json").read_text())
    e = np.loadtxt(GOLD / "test_base.txt", dtype=np.int64, comments="#")
    g1 = oracle.Graph.from_edges(32, e[:, 0], e[:, 1])
    off, col, _ = g1.csr()
    d, p, _ = run_sssp(gpu, off, col, None, 0)
    assert d.tolist() == exp["sssp_unit_dist"]
    w = (1 + (7 * e[:, 0] + 3 * e[:, 1]) % 10).astype(np.uint32)
    gw = oracle.Graph.from_edges(32, e[:, 0], e[:, 1], w)
    off, col, ww = gw.csr()
    d, p, _ = run_sssp(gpu, off, col, ww, 0)
    assert d.tolist() == exp["sssp_weighted_dist"]
    od, op = gw.dijkstra(0)
    assert np.array_equal(p, op)


@pytest.mark.parametrize("scale,src,wmax", [(10, 0, 64), (14, 3, 64), (16, 0, 255), (16, 12345, 1)])
def test_sssp_bit_exact_vs_dijkstra(gpu, scale, src, wmax):
    g = oracle.Graph.rmat(scale, 16, seed=2, weights="u32", wmin=1, wmax=wmax)
    off, col, w = g.csr()
    d, p, st = run_sssp(gpu, off, col, w, src)
    od, op = g.dijkstra(src)
    assert np.array_equal(d, od)
    assert np.array_equal(p, op)
    assert st["edges_traversed"] > 0


def test_sssp_heavy_vertex_chunks(gpu):
    """A star hub with out-degree 20000 > the 4096-edge chunk threshold."""
    n = 30000
    src = np.concatenate([np.zeros(20000, np.int64), np.arange(1, 20000)])
    dst = np.concatenate([np.arange(1, 20001), np.arange(2, 20001)])
    w = (1 + (src * 7 + dst) % 50).astype(np.uint32)
    g = oracle.Graph.from_edges(n, src, dst, w)
    off, col, ww = g.csr()
    d, p, _ = run_sssp(gpu, off, col, ww, 0)
    od, op = g.dijkstra(0)
    assert np.array_equal(d, od) and np.array_equal(p, op)


@pytest.mark.parametrize("frac,wmax", [(0.001, 64), (0.01, 64), (0.01, 255)])
def test_incremental_sssp_bit_exact(gpu, frac, wmax):
    g = oracle.Graph.rmat(14, 16, seed=2, weights="u32", wmin=1, wmax=wmax)
    off, col, w = g.csr()
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col, w)
        ctx.sssp_init(0)
        ctx.sssp_run()
        for b in range(4):
            n_ops = max(2, int(frac * g.m))
            ops = g.gen_batch(seed=2, batch_id=b, n_del=n_ops // 2, n_ins=n_ops - n_ops // 2, wmin=1, wmax=wmax)
            if b == 2:  # re-weight existing edges (up and down) through adds
                o2, c2, w2 = g.csr()
                extra = np.zeros(200, ops.dtype)
                us = np.searchsorted(o2, np.arange(0, g.m, g.m // 200)[:200], side="right") - 1
                extra["u"] = us
                extra["v"] = c2[np.arange(0, g.m, g.m // 200)[:200]]
                extra["w_u32"] = np.where(np.arange(200) % 2, 1, wmax)
                ops = np.concatenate([ops, extra])
            ctx.apply_batch(ops)
            g, _ = g.apply_batch(ops)
            ctx.sssp_run()
            d, p = ctx.sssp_download()
            od, op = g.dijkstra(0)
            assert np.array_equal(d, od), b
            assert np.array_equal(p, op), b


def test_apply_batch_before_first_run(gpu):
    g = oracle.Graph.rmat(12, 16, seed=2, weights="u32")
    off, col, w = g.csr()
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col, w)
        ctx.sssp_init(7)
        ops = g.gen_batch(seed=9, batch_id=0, n_del=100, n_ins=100)
        ctx.apply_batch(ops)
        g, _ = g.apply_batch(ops)
        ctx.sssp_run()
        d, p = ctx.sssp_download()
    od, op = g.dijkstra(7)
    assert np.array_equal(d, od) and np.array_equal(p, op)
