"""Pins the oracle (oracle/oracle.cpp) against known answers, closed forms and
brute force, so that it can serve as the checker of the HIP path.  CPU only."""
import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg
import scipy.sparse.linalg as sla

import oracle
from tests.conftest import DAMPING

GOLD = Path(__file__).resolve().parent / "golden"


# ---------------------------------------------------------------- Philox
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    assert list(oracle.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                                       0x6D5451FD]
    assert list(oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


# ---------------------------------------------------------------- R-MAT
STREAM_RMAT = 0x524D4154


def rmat_bruteforce(scale, ef, a, b, c, seed):
    """Independent restatement of the generator with Python ints."""
    t1 = int(a * 4294967296.0)
    t2 = int((a + b) * 4294967296.0)
    t3 = int((a + b + c) * 4294967296.0)
    key = [seed & 0xFFFFFFFF, seed >> 32]
    edges = set()
    for i in range(ef << scale):
        u = v = 0
        words = []
        for blk in range((scale + 3) // 4):
            words.extend(int(x) for x in oracle.philox([i & 0xFFFFFFFF, i >> 32, blk, STREAM_RMAT], key))
        for l in range(scale):
            r = words[l]
            q = 0 if r < t1 else 1 if r < t2 else 2 if r < t3 else 3
            u |= (q >> 1) << (scale - 1 - l)
            v |= (q & 1) << (scale - 1 - l)
        if u != v:
            edges.add((u, v))
    return sorted(edges)


def test_rmat_matches_bruteforce():
    g = oracle.Graph.rmat(5, 4, 0.57, 0.19, 0.19, seed=7)
    off, col, _ = g.csr()
    got = [(u, int(v)) for u in range(g.n) for v in col[off[u]:off[u + 1]]]
    assert got == rmat_bruteforce(5, 4, 0.57, 0.19, 0.19, 7)


def test_rmat_invariants_and_quadrants():
    s = 12
    g = oracle.Graph.rmat(s, 16, 0.57, 0.19, 0.19, seed=1)
    off, col, _ = g.csr()
    src = np.repeat(np.arange(g.n), np.diff(off).astype(np.int64))
    assert np.all(src != col)                              # no self-loops
    key = src.astype(np.int64) << 32 | col
    assert np.all(np.diff(key) > 0)                        # sorted, no duplicates
    assert 0.75 * (16 << s) < g.m < (16 << s)              # ~81% unique at scale 12
    # top-level quadrant frequencies of the unique edges stay near a,b,c,d
    top_u = src >> (s - 1)
    top_v = col.astype(np.int64) >> (s - 1)
    frac_a = np.mean((top_u == 0) & (top_v == 0))
    assert 0.5 < frac_a < 0.6


def test_rmat_weights_are_pure_function_of_pair():
    g1 = oracle.Graph.rmat(8, 8, seed=3, weights="u32", wmin=1, wmax=64)
    off, col, w = g1.csr()
    assert w.min() >= 1 and w.max() <= 64
    g2 = oracle.Graph.rmat(8, 16, seed=3, weights="u32", wmin=1, wmax=64)
    off2, col2, w2 = g2.csr()
    d1 = {(u, int(v)): int(x) for u in range(g1.n) for v, x in zip(col[off[u]:off[u + 1]], w[off[u]:off[u + 1]])}
    d2 = {(u, int(v)): int(x) for u in range(g2.n) for v, x in zip(col2[off2[u]:off2[u + 1]], w2[off2[u]:off2[u + 1]])}
    common = set(d1) & set(d2)
    assert len(common) > 100 and all(d1[k] == d2[k] for k in common)
    gf = oracle.Graph.rmat(8, 8, seed=3, weights="f32")
    wf = gf.csr()[2]
    assert wf.dtype == np.float32 and wf.min() > 0 and wf.max() <= 1


# ---------------------------------------------------------------- batches
def apply_bruteforce(edges: dict, ops):
    added, deleted = {}, set()
    for op in ops:
        k = (int(op["u"]), int(op["v"]))
        if op["type"] == 0:
            added.setdefault(k, int(op["w_u32"]))
        else:
            deleted.add(k)
    out = dict(added)
    for k, w in edges.items():
        if k not in deleted and k not in added:
            out[k] = w
    return out


def test_apply_batch_semantics_bruteforce():
    rng = np.random.default_rng(0)
    n = 40
    src = rng.integers(0, n, 300)
    dst = rng.integers(0, n, 300)
    keep = src != dst
    pairs = sorted(set(zip(src[keep].tolist(), dst[keep].tolist())))
    w = {p: int(rng.integers(1, 9)) for p in pairs}
    g = oracle.Graph.from_edges(n, [p[0] for p in pairs], [p[1] for p in pairs], np.array([w[p] for p in pairs],
                                                                                           np.uint32))
    ops = np.zeros(120, oracle.UPDATE_DTYPE)
    for i in range(120):
        if i % 3 == 0 and pairs:  # delete existing
            p = pairs[rng.integers(len(pairs))]
            ops[i] = (1, p[0], p[1], 0, 0, 0)
        elif i % 3 == 1:          # add (maybe existing, maybe duplicate)
            u, v = int(rng.integers(n)), int(rng.integers(n))
            ops[i] = (0, u, v if v != u else (v + 1) % n, int(rng.integers(1, 9)), 0, 0)
        else:                     # delete a random (maybe absent) pair
            u, v = int(rng.integers(n)), int(rng.integers(n))
            ops[i] = (1, u, v, 0, 0, 0)
    ops[5] = (0, ops[3]["u"], ops[3]["v"], 99, 0, 0)   # add + delete of the same pair: add wins
    ops[7] = (0, ops[4]["u"], ops[4]["v"], 77, 0, 0)   # second add of a pair: first add wins
    ng, touched = g.apply_batch(ops)
    off, col, wu = ng.csr()
    got = {(u, int(v)): int(x) for u in range(n) for v, x in zip(col[off[u]:off[u + 1]], wu[off[u]:off[u + 1]])}
    assert got == apply_bruteforce(w, ops)
    assert set(np.nonzero(touched)[0].tolist()) == set(int(x) for x in ops["u"])


def test_gen_batch_shape():
    g = oracle.Graph.rmat(10, 16, seed=2, weights="u32")
    ops = g.gen_batch(seed=2, batch_id=0, n_del=50, n_ins=60, wmin=1, wmax=64)
    off, col, _ = g.csr()
    dels, ins = ops[:50], ops[50:]
    assert np.all(dels["type"] == 1) and np.all(ins["type"] == 0)
    for op in dels:  # deletes name existing edges
        assert op["v"] in col[off[op["u"]]:off[op["u"] + 1]]
    assert np.all(ins["u"] != ins["v"])
    assert ins["w_u32"].min() >= 1 and ins["w_u32"].max() <= 64


# ---------------------------------------------------------------- PageRank
def closed_form(g, d):
    off, col, _ = g.csr()
    deg = np.diff(off).astype(np.int64)
    src = np.repeat(np.arange(g.n), deg)
    P = sp.csr_matrix((1.0 / deg[src], (src, col.astype(np.int64))), shape=(g.n, g.n))
    return sla.spsolve((sp.identity(g.n) - d * P.T).tocsc(), np.full(g.n, (1 - d) / g.n))


def test_pagerank_golden_test_base():
    exp = json.loads((GOLD / "test_base_expected.json").read_text())
    e = np.loadtxt(GOLD / "test_base.txt", dtype=np.int64, comments="#")
    g = oracle.Graph.from_edges(32, e[:, 0], e[:, 1])
    r, delta, st = g.pr_solve(exp["damping"], tol=1e-15)
    assert np.abs(r - np.array(exp["pagerank"])).max() < 1e-13
    # the reference's own stopping rule (per-vertex T/N, sum|delta| < T)
    r2, delta2, st2 = g.pr_ingress(exp["damping"], T=1e-10 * 32, max_rounds=1000)
    assert np.abs(r2 - np.array(exp["pagerank"])).sum() <= st2["residual"] / (1 - exp["damping"]) + 1e-15


@pytest.mark.parametrize("scale", [10, 13])
def test_pagerank_ingress_vs_closed_form(scale):
    g = oracle.Graph.rmat(scale, 16, seed=1)
    rs = closed_form(g, DAMPING)
    T = 1e-9 * g.n
    r, delta, st = g.pr_ingress(DAMPING, T, max_rounds=1000)
    err = np.abs(r - rs).sum()
    bound = st["residual"] / (1 - DAMPING)
    assert st["residual"] < T
    assert err <= bound * (1 + 1e-9) + 1e-12
    # mass invariant: sum r + sum delta - d * sum_{outdeg>0} r = 1 - d
    off, _, _ = g.csr()
    nd = np.diff(off) > 0
    assert abs(r.sum() + delta.sum() - DAMPING * r[nd].sum() - (1 - DAMPING)) < 1e-12
    rt, _, _ = g.pr_solve(DAMPING, tol=1e-14)
    assert np.abs(rt - rs).max() < 1e-12


def test_pagerank_check_detects_planted_dangling_mistake():
    """A teleporting (normalised) PageRank differs from the reference's leaky one
    by far more than the tolerance the parity tests use."""
    g = oracle.Graph.rmat(10, 16, seed=1)
    rs = closed_form(g, DAMPING)
    off, col, _ = g.csr()
    deg = np.diff(off).astype(np.int64)
    src = np.repeat(np.arange(g.n), deg)
    P = sp.csr_matrix((1.0 / deg[src], (src, col.astype(np.int64))), shape=(g.n, g.n)).toarray()
    P[deg == 0, :] = 1.0 / g.n
    wrong = np.linalg.solve(np.eye(g.n) - DAMPING * P.T, np.full(g.n, (1 - DAMPING) / g.n))
    assert np.abs(wrong - rs).sum() > 1e-3


def test_amend_reseed_reaches_new_fixed_point():
    g = oracle.Graph.rmat(11, 16, seed=3)
    r, delta, _ = g.pr_solve(DAMPING, tol=1e-15)
    ops = g.gen_batch(seed=3, batch_id=0, n_del=300, n_ins=300)
    ng, touched = g.apply_batch(ops)
    for use_touched in (True, False):
        rr, dd = r.copy(), delta.copy()
        g.pr_amend(DAMPING, -1, rr, dd, touched if use_touched else None)
        ng.pr_amend(DAMPING, +1, rr, dd, touched if use_touched else None)
        r2, d2, _ = ng.pr_solve(DAMPING, tol=1e-15, r=rr, delta=dd)
        assert np.abs(r2 - closed_form(ng, DAMPING)).max() < 1e-12


# ---------------------------------------------------------------- SSSP
def scipy_dist(g, src):
    off, col, w = g.csr()
    deg = np.diff(off).astype(np.int64)
    s = np.repeat(np.arange(g.n), deg)
    ww = np.ones(g.m) if w is None else w.astype(float)
    G = sp.csr_matrix((ww, (s, col.astype(np.int64))), shape=(g.n, g.n))
    d = csg.dijkstra(G, indices=src)
    return np.where(np.isinf(d), 0xFFFFFFFF, d).astype(np.uint64)


def check_parents(g, src, dist, par):
    off, col, w = g.csr()
    INF = 0xFFFFFFFF
    best = np.full(g.n, INF, np.int64)
    for u in range(g.n):
        if dist[u] == INF:
            continue
        for k in range(int(off[u]), int(off[u + 1])):
            v = int(col[k])
            if v != src and int(dist[u]) + (1 if w is None else int(w[k])) == int(dist[v]):
                best[v] = min(best[v], u)
    for v in range(g.n):
        if v == src or dist[v] == INF:
            assert par[v] == v
        else:
            assert par[v] == best[v]


def test_sssp_golden_test_base():
    exp = json.loads((GOLD / "test_base_expected.json").read_text())
    e = np.loadtxt(GOLD / "test_base.txt", dtype=np.int64, comments="#")
    g1 = oracle.Graph.from_edges(32, e[:, 0], e[:, 1])
    d1, p1 = g1.dijkstra(0)
    assert d1.tolist() == exp["sssp_unit_dist"]
    w = (1 + (7 * e[:, 0] + 3 * e[:, 1]) % 10).astype(np.uint32)
    gw = oracle.Graph.from_edges(32, e[:, 0], e[:, 1], w)
    dw, pw = gw.dijkstra(0)
    assert dw.tolist() == exp["sssp_weighted_dist"]
    check_parents(gw, 0, dw, pw)
    di, pi, _ = gw.sssp_ingress(0)
    assert np.array_equal(di, dw) and np.array_equal(pi, pw)


@pytest.mark.parametrize("scale,src", [(9, 0), (11, 5)])
def test_dijkstra_and_ingress_vs_scipy(scale, src):
    g = oracle.Graph.rmat(scale, 16, seed=2, weights="u32", wmin=1, wmax=64)
    dist, par = g.dijkstra(src)
    assert np.array_equal(dist.astype(np.uint64), scipy_dist(g, src))
    check_parents(g, src, dist, par)
    di, pi, st = g.sssp_ingress(src)
    assert np.array_equal(di, dist) and np.array_equal(pi, par)
    assert st["edges"] > 0


@pytest.mark.parametrize("weights", [None, "u32"])
def test_sssp_incremental_step_equals_dijkstra(weights):
    """The restated deltaCompute + resend + Ingress step reaches the new graph's
    Dijkstra result (distances and smallest-id parents) batch after batch,
    including re-weighted parent edges."""
    g = oracle.Graph.rmat(12, 16, seed=6, weights=weights, wmin=1, wmax=64)
    src = 0
    dist, par = g.dijkstra(src)
    for b in range(4):
        ops = g.gen_batch(seed=6, batch_id=b, n_del=200, n_ins=200)
        if weights and b == 1:
            # re-weight ten parent edges upward: their children must be reset
            off, col, w = g.csr()
            kids = [v for v in range(g.n) if dist[v] not in (0, 0xFFFFFFFF)][:10]
            extra = np.zeros(len(kids), ops.dtype)
            for i, v in enumerate(kids):
                extra[i] = (0, par[v], v, 200, 0.0, 0)
            ops = np.concatenate([ops, extra])
        g2, _ = g.apply_batch(ops)
        d2, p2, st = g2.sssp_inc(g, src, dist, par)
        od, op = g2.dijkstra(src)
        assert np.array_equal(d2, od), b
        assert np.array_equal(p2, op), b
        g, dist, par = g2, d2, p2
