"""Writes the expected outputs for the golden inputs in this folder.

Expected values come from closed forms, not from any implementation under test:
  * PageRank: the fixed point of the reference's delta-PageRank
    (pagerank_ingress.h:27-70: delta0 = (1-d)/N, push d*delta/outdeg, no push from
    dangling vertices) solves (I - d P^T) r = (1-d)/N * 1 with P row-stochastic on
    non-dangling rows and zero on dangling rows; solved with scipy's sparse LU.
  * SSSP: scipy.sparse.csgraph.dijkstra from source 0 with unit weights and with
    the deterministic weights w(u,v) = 1 + (7u + 3v) mod 10.
Run: python tests/golden/make_golden.py
"""
import json
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csg
import scipy.sparse.linalg as sla

HERE = Path(__file__).resolve().parent
D = float(np.float32(0.85))


def load_edges():
    e = [tuple(map(int, l.split())) for l in (HERE / "test_base.txt").read_text().splitlines()
         if l.strip() and not l.startswith("#")]
    return np.array(e, dtype=np.int64)


def pagerank_closed_form(n, src, dst, d):
    deg = np.bincount(src, minlength=n).astype(float)
    P = sp.csr_matrix((1.0 / deg[src], (src, dst)), shape=(n, n))
    A = (sp.identity(n) - d * P.T).tocsc()
    return sla.spsolve(A, np.full(n, (1 - d) / n))


def main():
    e = load_edges()
    n = 32
    src, dst = e[:, 0], e[:, 1]
    r = pagerank_closed_form(n, src, dst, D)
    w = 1 + (7 * src + 3 * dst) % 10
    G1 = sp.csr_matrix((np.ones(len(src)), (src, dst)), shape=(n, n))
    Gw = sp.csr_matrix((w.astype(float), (src, dst)), shape=(n, n))
    d1 = csg.dijkstra(G1, indices=0)
    dw = csg.dijkstra(Gw, indices=0)
    inf = 0xFFFFFFFF
    out = {
        "n": n,
        "m": int(len(src)),
        "damping": D,
        "pagerank": r.tolist(),
        "sssp_unit_dist": [inf if np.isinf(x) else int(x) for x in d1],
        "sssp_weight_rule": "1 + (7*u + 3*v) % 10",
        "sssp_weighted_dist": [inf if np.isinf(x) else int(x) for x in dw],
    }
    (HERE / "test_base_expected.json").write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", HERE / "test_base_expected.json")


if __name__ == "__main__":
    main()
