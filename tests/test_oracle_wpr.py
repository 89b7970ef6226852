"""Pins the oracle's weighted personalized PageRank and its adjoint: closed-form
sparse solves of both linear systems and central finite differences of the FP64
forward for the edge-weight and personalization gradients.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle
from tests.conftest import DAMPING


def setup(scale=10, seeds=32, seed=4):
    g = oracle.Graph.rmat(scale, 8, seed=seed, weights="f32")
    off, col, w = g.csr()
    rng = np.random.default_rng(seed)
    p = np.zeros(g.n)
    idx = rng.choice(g.n, seeds, replace=False)
    p[idx] = rng.dirichlet(np.ones(seeds))
    c = rng.standard_normal(g.n)
    return g, off, col, w, p, c, rng


def row_normalised(g, off, col, w):
    deg = np.diff(off).astype(np.int64)
    src = np.repeat(np.arange(g.n), deg)
    s = np.bincount(src, weights=w.astype(np.float64), minlength=g.n)
    return sp.csr_matrix((w.astype(np.float64) / s[src], (src, col.astype(np.int64))), shape=(g.n, g.n))


def test_forward_and_adjoint_equal_closed_forms():
    g, off, col, w, p, c, _ = setup()
    W = row_normalised(g, off, col, w)
    r, _ = g.wpr_forward(w, p, DAMPING)
    rs = sla.spsolve((sp.identity(g.n) - DAMPING * W.T).tocsc(), (1 - DAMPING) * p)
    assert np.abs(r - rs).max() < 1e-14
    lam, gp, gw = g.wpr_adjoint(w, r, c, DAMPING)
    ls = sla.spsolve((sp.identity(g.n) - DAMPING * W).tocsc(), c)
    assert np.abs(lam - ls).max() < 1e-12
    assert np.allclose(gp, (1 - DAMPING) * lam, rtol=0, atol=0)
    # dangling rows (R-MAT has many) carry no weight gradient and leak their mass
    assert (np.diff(off) == 0).any()


def test_weight_gradient_equals_central_differences():
    g, off, col, w, p, c, rng = setup()
    r, _ = g.wpr_forward(w, p, DAMPING)
    _, _, gw = g.wpr_adjoint(w, r, c, DAMPING)

    def loss(wv):
        return c @ g.wpr_forward(wv, p, DAMPING)[0]

    for e in rng.choice(len(w), 24, replace=False):
        w1, w2 = w.copy(), w.copy()
        w1[e] += np.float32(1e-3)
        w2[e] -= np.float32(1e-3)
        fd = (loss(w1) - loss(w2)) / (np.float64(w1[e]) - np.float64(w2[e]))
        assert abs(fd - gw[e]) <= 1e-5 * abs(fd) + 1e-12, (e, fd, gw[e])


def test_personalization_gradient_equals_central_differences():
    g, off, col, w, p, c, rng = setup()
    r, _ = g.wpr_forward(w, p, DAMPING)
    _, gp, _ = g.wpr_adjoint(w, r, c, DAMPING)
    for v in rng.choice(g.n, 16, replace=False):
        p1, p2 = p.copy(), p.copy()
        p1[v] += 1e-3
        p2[v] -= 1e-3
        fd = (c @ g.wpr_forward(w, p1, DAMPING)[0] - c @ g.wpr_forward(w, p2, DAMPING)[0]) / 2e-3
        assert abs(fd - gp[v]) <= 1e-8 * max(1.0, abs(fd)), (v, fd, gp[v])


def test_fd_check_detects_a_dropped_normalisation_term():
    """A gradient missing the row-normalisation term (dL/dw = d r_u/s_u lambda_v)
    fails the finite-difference check: the check pins the formula."""
    g, off, col, w, p, c, rng = setup()
    r, _ = g.wpr_forward(w, p, DAMPING)
    lam, _, gw = g.wpr_adjoint(w, r, c, DAMPING)
    deg = np.diff(off).astype(np.int64)
    src = np.repeat(np.arange(g.n), deg)
    s = np.bincount(src, weights=w.astype(np.float64), minlength=g.n)
    wrong = DAMPING * r[src] / s[src] * lam[col]
    bad = 0
    for e in rng.choice(len(w), 24, replace=False):
        w1, w2 = w.copy(), w.copy()
        w1[e] += np.float32(1e-3)
        w2[e] -= np.float32(1e-3)
        fd = (c @ g.wpr_forward(w1, p, DAMPING)[0] - c @ g.wpr_forward(w2, p, DAMPING)[0]) / (
            np.float64(w1[e]) - np.float64(w2[e]))
        bad += abs(fd - wrong[e]) > 1e-5 * abs(fd) + 1e-12
    assert bad > 12
