"""The oracle's BFS, SSWP and PHP restatements, pinned without a GPU: BFS against
scipy's unweighted shortest paths, SSWP against brute force over every simple
path, PHP against closed forms worked by hand and the explicit series."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import shortest_path

import oracle
from oracle import apps

INF = 0xFFFFFFFF
D = float(np.float32(0.8))


def _csr(n, edges, weights=None):
    edges = sorted(edges)
    off = np.zeros(n + 1, np.int64)
    for u, _ in edges:
        off[u + 1] += 1
    off = np.cumsum(off)
    col = np.array([v for _, v in edges], np.int64)
    w = None if weights is None else np.array([weights[e] for e in edges], np.uint32)
    return off, col, w


def test_bfs_equals_scipy_unweighted():
    g = oracle.Graph.rmat(10, 8, seed=3)
    off, col, _ = g.csr()
    A = sp.csr_matrix((np.ones(len(col)), col.astype(np.int64), off.astype(np.int64)), shape=(g.n, g.n))
    ref = shortest_path(A, unweighted=True, indices=0)
    got = apps.bfs(off, col, 0)
    assert np.array_equal(np.where(np.isinf(ref), INF, ref).astype(np.uint32), got)


@pytest.mark.parametrize("seed", range(6))
def test_sswp_equals_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = 7
    edges = {(int(u), int(v)) for u, v in rng.integers(0, n, (16, 2)) if u != v}
    weights = {e: int(rng.integers(1, 20)) for e in edges}
    off, col, w = _csr(n, edges, weights)
    got = apps.sswp(off, col, w, 0)
    best = np.zeros(n, np.uint64)
    best[0] = INF
    for k in range(1, n):
        for path in itertools.permutations(range(1, n), k):
            p = (0,) + path
            if all(e in edges for e in zip(p, p[1:])):
                width = min(weights[e] for e in zip(p, p[1:]))
                best[p[-1]] = max(best[p[-1]], width)
    assert np.array_equal(got, best.astype(np.uint32))


def test_php_closed_forms():
    # chain s=0 -> 1 -> 2 with 1 -> 0 as well: 1 sends only to 2 (edges into s are
    # dropped and left out of s_1): values 1, d, d^2
    off, col, w = _csr(3, [(0, 1), (1, 2), (1, 0)], {(0, 1): 3, (1, 2): 5, (1, 0): 7})
    assert np.allclose(apps.php(off, col, w, 0, D), [1, D, D * D], rtol=1e-14)
    # two-cycle behind the source: s -> a, a -> b, b -> a: a = d / (1 - d^2), b = d^2 / (1 - d^2)
    off, col, _ = _csr(3, [(0, 1), (1, 2), (2, 1)])
    assert np.allclose(apps.php(off, col, None, 0, D), [1, D / (1 - D * D), D * D / (1 - D * D)], rtol=1e-14)
    # weights split the mass: s -> a (w 1), s -> b (w 3)
    off, col, w = _csr(3, [(0, 1), (0, 2)], {(0, 1): 1, (0, 2): 3})
    assert np.allclose(apps.php(off, col, w, 0, D), [1, D / 4, 3 * D / 4], rtol=1e-14)


def test_php_equals_the_series():
    g = oracle.Graph.rmat(9, 8, seed=1, weights="u32", wmin=1, wmax=9)
    off, col, w = g.csr()
    src = 3
    n = g.n
    x = np.zeros(n)
    x[src] = 1.0
    val = np.zeros(n)
    rows = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    c64 = col.astype(np.int64)
    wf = w.astype(np.float64)
    keep = c64 != src
    s = np.bincount(rows[keep], weights=wf[keep], minlength=n)
    for _ in range(400):
        val += x
        nx = np.zeros(n)
        np.add.at(nx, c64[keep], D * wf[keep] / s[rows[keep]] * x[rows[keep]])
        x = nx
    assert np.abs(x).sum() < 1e-20
    assert np.allclose(apps.php(off, col, w, src, D), val, rtol=1e-10, atol=1e-15)
