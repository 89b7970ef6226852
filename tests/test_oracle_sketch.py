"""Pins the oracle's restatement of the supernode compressor and the SumInc
PageRank query: hand-computed structure, dense closed-form shortcut rows and the
closed-form fixed point.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle
from tests.conftest import DAMPING

INF = 0xFFFFFFFF


def closed_form(g, d):
    off, col, _ = g.csr()
    deg = np.diff(off).astype(np.int64)
    src = np.repeat(np.arange(g.n), deg)
    P = sp.csr_matrix((1.0 / deg[src], (src, col.astype(np.int64))), shape=(g.n, g.n))
    return sla.spsolve((sp.identity(g.n) - d * P.T).tocsc(), np.full(g.n, (1 - d) / g.n))


def test_leaf_clusters_hand_example():
    # hub 0 with private leaves 1,2,3 (0->1, 0->2, 3->0); 4 (neighbours 0 and 7) is
    # no leaf but anchors its private leaf 7; leaf pair 5<->6 anchors on 5; 8 isolated.
    # Ids follow anchor order: 0 -> 0, 4 -> 1, 5 -> 2.
    e = [(0, 1), (0, 2), (3, 0), (0, 4), (4, 7), (7, 4), (5, 6), (6, 5)]
    g = oracle.Graph.from_edges(9, [a for a, b in e], [b for a, b in e])
    raw = g.leaf_clusters(max_node=1024, min_node=2)
    assert raw.tolist() == [0, 0, 0, 0, 1, 2, 2, 1, INF]
    raw3 = g.leaf_clusters(max_node=3, min_node=2)  # hub keeps its first 2 leaves
    assert raw3.tolist() == [0, 0, 0, INF, 1, 2, 2, 1, INF]
    raw_min4 = g.leaf_clusters(max_node=1024, min_node=4)
    assert raw_min4.tolist() == [0, 0, 0, 0, INF, INF, INF, INF, INF]


def test_structure_hand_example():
    # cluster A = {0,1,2,3}: 9->0 enters, 2->9 leaves; internal 0->1,1->2,2->3,3->0,0->2 (5 edges)
    #   entries {0}, exits {2}: benefit 5 - 1*1 = 4 > 1 -> accepted; 0 OnlyIn, 2 OnlyOut, 1,3 Inner
    # cluster B = {4,5}: 4->5 internal (1 edge), 9->4, 5->9, 9->5, 4->9: entries {4,5}, exits {4,5}
    #   benefit 1 - 2*2 < 1 -> rejected -> Singles
    e = [(9, 0), (2, 9), (0, 1), (1, 2), (2, 3), (3, 0), (0, 2), (4, 5), (9, 4), (5, 9), (9, 5), (4, 9)]
    g = oracle.Graph.from_edges(10, [a for a, b in e], [b for a, b in e])
    raw = np.array([0, 0, 0, 0, 1, 1, INF, INF, INF, INF], np.uint32)
    cid, typ = g.sum_structure(raw, min_node=2, max_node=1024, threshold=1.0)
    assert cid.tolist() == [0, 0, 0, 0, INF, INF, INF, INF, INF, INF]
    assert typ.tolist() == [2, 4, 1, 4, 0, 0, 0, 0, 0, 0]
    # forcing acceptance shows the Both type on B
    cid2, typ2 = g.sum_structure(raw, min_node=2, max_node=1024, threshold=-100.0)
    assert cid2.tolist()[:6] == [0, 0, 0, 0, 1, 1] and typ2.tolist()[4:6] == [3, 3]


def shortcut_dense(g, cid, entry, d):
    """Row `entry` of (I - d A_C)^-1 with A_C[u,v] = 1/outdeg(u) on edges inside the cluster."""
    off, col, _ = g.csr()
    k = cid[entry]
    mem = np.nonzero(cid == k)[0]
    pos = {int(v): i for i, v in enumerate(mem)}
    A = np.zeros((len(mem), len(mem)))
    for v in mem:
        deg = off[v + 1] - off[v]
        for x in col[off[v]:off[v + 1]]:
            if cid[x] == k:
                A[pos[int(v)], pos[int(x)]] += 1.0 / deg
    e = np.zeros(len(mem))
    e[pos[int(entry)]] = 1
    return mem, np.linalg.solve((np.eye(len(mem)) - d * A).T, e)


def test_shortcut_rows_equal_dense_solve():
    g = oracle.Graph.rmat(12, 16, seed=1)
    raw = g.leaf_clusters(1024, 2)
    cid, typ = g.sum_structure(raw, 2, 1024, 1.0)
    entries = np.nonzero((typ == 2) | (typ == 3))[0]
    assert len(entries) > 10
    for s in entries[:10]:
        r, dl, steps = g.sum_shortcut(cid, int(s), DAMPING, thr=1e-16, maxit=2000)
        mem, x = shortcut_dense(g, cid, int(s), DAMPING)
        assert np.abs(r[mem] - x).max() < 1e-12
        assert np.abs(dl).sum() < 1e-12
        outside = np.ones(g.n, bool)
        outside[mem] = False
        assert not r[outside].any()


def test_shortcut_truncation_keeps_mass():
    """A truncated shortcut still accounts for the unit: the leftover delta plus the
    mass that left the cluster or evaporated equals the injection."""
    g = oracle.Graph.rmat(11, 16, seed=4)
    raw = np.where(np.arange(g.n) < 256, np.arange(g.n) // 16, INF).astype(np.uint32)
    cid, typ = g.sum_structure(raw, 2, 1024, -1e9)
    off, col, _ = g.csr()
    s = int(np.nonzero((typ == 2) | (typ == 3))[0][0])
    r, dl, steps = g.sum_shortcut(cid, s, DAMPING, thr=0.0, maxit=5)
    k = cid[s]
    # mass balance: injected 1 = (1-d) sum r + d * sum_v r_v * (external out-share + dangling 0) + leftover
    ext = 0.0
    for v in np.nonzero(cid == k)[0]:
        deg = off[v + 1] - off[v]
        if deg:
            ext += DAMPING * r[v] * np.count_nonzero(cid[col[off[v]:off[v + 1]]] != k) / deg
        else:
            ext += DAMPING * r[v]  # a dangling member leaks its d share
    assert abs((1 - DAMPING) * r.sum() + ext + dl.sum() - 1.0) < 1e-12


@pytest.mark.parametrize("cluster_kind", ["leaf", "blocks"])
def test_sum_pr_equals_closed_form(cluster_kind):
    g = oracle.Graph.rmat(12, 16, seed=1)
    if cluster_kind == "leaf":
        raw = g.leaf_clusters(1024, 2)
        thr = 1.0
    else:
        raw = np.where(np.arange(g.n) < 2048, np.arange(g.n) // 8, INF).astype(np.uint32)
        thr = -1e9  # accept every block so all node types occur
    cid, typ = g.sum_structure(raw, 2, 1024, thr)
    if cluster_kind == "blocks":
        assert set(np.unique(typ).tolist()) == {0, 1, 2, 3, 4}
    T = 1e-10 * g.n
    r, dl, st = g.sum_pr(cid, typ, DAMPING, T)
    rs = closed_form(g, DAMPING)
    assert st["sketch_rounds"] > 0
    assert np.abs(r - rs).sum() <= st["residual"] / (1 - DAMPING) + 1e-12


def _weighted(g):
    off, col, w = g.csr()
    return off, col, (np.ones(len(col)) if w is None else w.astype(np.float64))


def test_sum_sssp_shortcut_equals_induced_dijkstra():
    from scipy.sparse.csgraph import dijkstra
    g = oracle.Graph.rmat(12, 16, seed=1, weights="u32", wmin=1, wmax=64)
    raw = np.where(np.arange(g.n) < 2048, np.arange(g.n) // 16, INF).astype(np.uint32)
    cid, typ = g.sum_structure(raw, 2, 1024, -1e9)
    off, col, w = _weighted(g)
    entries = np.nonzero((typ == 2) | (typ == 3))[0]
    assert len(entries) > 20
    for s in entries[:20]:
        mem = np.nonzero(cid == cid[s])[0]
        pos = {int(v): i for i, v in enumerate(mem)}
        A = np.zeros((len(mem), len(mem)))
        for v in mem:
            for e in range(off[v], off[v + 1]):
                if cid[col[e]] == cid[s]:
                    A[pos[int(v)], pos[int(col[e])]] = w[e]
        ref = dijkstra(sp.csr_matrix(A), indices=pos[int(s)])
        got = g.sum_sssp_shortcut(cid, int(s))
        exp = np.full(g.n, INF, np.uint32)
        fin = np.isfinite(ref)
        exp[mem[fin]] = ref[fin].astype(np.uint32)
        assert np.array_equal(got, exp)


@pytest.mark.parametrize("weights", [None, "u32"])
@pytest.mark.parametrize("cluster_kind", ["leaf", "blocks"])
def test_sum_sssp_equals_dijkstra(weights, cluster_kind):
    g = oracle.Graph.rmat(12, 16, seed=2, weights=weights, wmin=1, wmax=64)
    if cluster_kind == "leaf":
        cid, typ = g.sum_structure(g.leaf_clusters(1024, 2), 2, 1024, 1.0)
    else:
        raw = np.where(np.arange(g.n) < 2048, np.arange(g.n) // 8, INF).astype(np.uint32)
        cid, typ = g.sum_structure(raw, 2, 1024, -1e9)
        assert set(np.unique(typ).tolist()) == {0, 1, 2, 3, 4}
    # a Single source, and a clustered source of every member type present
    sources = [int(np.nonzero(typ == 0)[0][0])]
    for t in (1, 2, 3, 4):
        if np.any(typ == t):
            sources.append(int(np.nonzero(typ == t)[0][0]))
    for s in sources:
        dist, par = g.dijkstra(s)
        d2, p2, st = g.sum_sssp(cid, typ, s)
        assert np.array_equal(d2, dist), (s, typ[s])
        assert np.array_equal(p2, par), (s, typ[s])
