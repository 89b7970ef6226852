"""CPU restatements of the reference's other Ingress apps run on the library's
kernels: BFS, SSWP (widest path) and PHP (penalized hitting probability).

TEST INFRASTRUCTURE ONLY: the checker of suminc_traversal_init (BFS, SSWP) and
suminc_php_run; nothing in the product imports it.

Pinned by tests/test_oracle_apps.py: BFS against scipy's unweighted shortest
paths; SSWP against brute force over every simple path of small graphs; PHP
against closed forms worked by hand (a chain, a two-cycle) and against the
explicit series on a small graph.
"""
from __future__ import annotations

import heapq
from collections import deque

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

INF = 0xFFFFFFFF


def bfs(off, col, src):
    """BFSIngress (bfs/bfs_ingress.h:60-95): hops from src, every edge counts 1;
    INF for unreached."""
    n = len(off) - 1
    dist = np.full(n, INF, np.uint64)
    dist[src] = 0
    q = deque([src])
    while q:
        u = q.popleft()
        for e in range(int(off[u]), int(off[u + 1])):
            v = int(col[e])
            if dist[v] == INF:
                dist[v] = dist[u] + 1
                q.append(v)
    return dist.astype(np.uint32)


def sswp(off, col, w, src):
    """SSWPIngress (sswp/sswp_ingress.h:60-140): width[v] = the largest over paths
    src -> v of the smallest edge weight on the path; width[src] = 0xFFFFFFFF (the
    numeric max it starts from), 0 for unreached (the identity element).  A
    Dijkstra on (max, min): the widest unsettled vertex is final."""
    n = len(off) - 1
    width = np.zeros(n, np.uint64)
    width[src] = INF
    done = np.zeros(n, bool)
    heap = [(-INF, src)]
    while heap:
        negw, u = heapq.heappop(heap)
        if done[u]:
            continue
        done[u] = True
        for e in range(int(off[u]), int(off[u + 1])):
            v = int(col[e])
            c = min(int(width[u]), int(w[e]))
            if c > width[v]:
                width[v] = c
                heapq.heappush(heap, (-c, v))
    return width.astype(np.uint32)


def php(off, col, w, src, d):
    """PHPIngress (php/php_ingress.h:15-120): value = sum over rounds of the deltas
    a vertex takes, from delta_0 = e_src, u sending d * w_uv / s_u of its delta to
    every out-neighbour v != src, s_u = the sum of u's out-weights to vertices other
    than src.  In closed form value = (I - A^T)^-1 e_src with A_uv = d w_uv / s_u
    (0 into src), solved in FP64."""
    n = len(off) - 1
    w = np.ones(len(col)) if w is None else np.asarray(w, np.float64)
    rows = np.repeat(np.arange(n), np.diff(np.asarray(off, np.int64)))
    col = np.asarray(col, np.int64)
    keep = col != src
    s = np.bincount(rows[keep], weights=w[keep], minlength=n)
    a = np.where(keep, d * w / np.where(s[rows] > 0, s[rows], 1.0), 0.0)
    A = sp.csr_matrix((a, (rows, col)), shape=(n, n))
    e = np.zeros(n)
    e[src] = 1.0
    return spla.spsolve((sp.identity(n, format="csc") - A.T.tocsc()), e)
