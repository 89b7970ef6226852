"""Graphs, batches and clusters read from the reference's text formats drive the
device path exactly as the same data passed as arrays.

  * an edge file uploads the CSR the host parser built (download_csr equal), and
    the golden 32-vertex file gives the closed-form PageRank;
  * an update file applied with apply_batch gives the graph the oracle's batch
    semantics give (set semantics: first add wins, add beats delete);
  * a cluster file written from the leaf clustering and read back builds the same
    supernode structure as the assignment itself.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.conftest import DAMPING

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def test_golden_edge_file_pagerank(gpu):
    from suminc_amd import textio

    exp = json.loads((GOLD / "test_base_expected.json").read_text())
    with textio.EdgeFile(GOLD / "test_base.txt") as ef, gpu.Context(pr_precision=gpu.PR_FP64,
                                                                   termcheck_threshold=1e-12) as ctx:
        ef.upload(ctx)
        off, col, _ = ef.csr()
        o2, c2, _ = ctx.download_csr()
        assert np.array_equal(off, o2) and np.array_equal(col, c2)
        ctx.pr_init()
        st = ctx.pr_run()
        r, _ = ctx.pr_download()
    err = np.abs(r - np.array(exp["pagerank"])).sum()
    assert err <= st["residual"] / (1 - DAMPING) + 1e-13


def test_weighted_edge_and_update_files(gpu, tmp_path):
    from suminc_amd import textio

    g = oracle.Graph.rmat(10, 8, seed=5, weights="u32", wmin=1, wmax=20)
    off, col, w = g.csr()
    rows = np.repeat(np.arange(g.n), np.diff(off).astype(np.int64))
    lines = [f"{u} {v} {x}" for u, v, x in zip(rows, col, w)]
    (tmp_path / "g_w.e").write_text("# weighted\n" + "\n".join(lines) + "\n")
    ef = textio.EdgeFile(tmp_path / "g_w.e", weights="u32")
    f_off, f_col, f_w = ef.csr()
    k = ef.n + 1  # the file fixes n as the largest id + 1
    assert np.array_equal(f_off, off[:k]) and np.all(off[k:] == off[k - 1])
    assert np.array_equal(f_col, col) and np.array_equal(f_w, w)
    ops = g.gen_batch(seed=5, batch_id=0, n_del=200, n_ins=200, wmin=1, wmax=20)
    ops = ops[(ops["u"] < ef.n) & (ops["v"] < ef.n)]
    ul = [f"a {o['u']} {o['v']} {o['w_u32']}" if int(o["type"]) == 0 else f"d {o['u']} {o['v']}" for o in ops]
    (tmp_path / "g_w.update").write_text("\n".join(ul) + "\n")
    g2, _ = oracle.Graph.from_csr(f_off, f_col, f_w).apply_batch(ops)
    o2, c2, w2 = g2.csr()
    with gpu.Context() as ctx:
        ef.upload(ctx)
        ef.close()
        ops_file = textio.read_updates(tmp_path / "g_w.update", weights="u32")
        assert len(ops_file) == len(ops)
        ctx.apply_batch(ops_file)
        d_off, d_col, d_w = ctx.download_csr()
    assert np.array_equal(d_off, o2) and np.array_equal(d_col, c2) and np.array_equal(d_w, w2)


def test_cluster_file_builds_the_same_sketch(gpu, tmp_path):
    from suminc_amd import textio

    g = oracle.Graph.rmat(12, 16, seed=2)
    off, col, _ = g.csr()
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col)
        cof = ctx.cluster_leaves(1024, 2)
        a = ctx.sum_build(2, 1024, 1.0)
        ta, _ = ctx.sum_download()
    path = tmp_path / textio.cluster_file_path("g_w.e", 1024)
    textio.write_clusters(path, cof)
    back, k = textio.read_clusters(path, g.n, min_node_num=2)
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col)
        ctx.set_clusters(back)
        b = ctx.sum_build(2, 1024, 1.0)
        tb, _ = ctx.sum_download()
    for key in ("clusters", "entries", "clustered_vertices", "internal_edges", "shortcut_slots", "node_type_count"):
        assert a[key] == b[key], key
    assert np.array_equal(ta != 0xFFFFFFFF, tb != 0xFFFFFFFF)
