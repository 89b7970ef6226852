"""CPU restatement of the library's community clustering (suminc_cluster_lpa), for
small graphs.

TEST INFRASTRUCTURE ONLY: the checker of suminc_cluster_lpa; nothing in the product
imports it.  The reference takes its clusters from an external Louvain pipeline
(sh/louvain_all.sh:24-49) whose membership is not pinned; this restates the
library's deterministic stand-in rule by rule, and tests/test_clusters.py pins it
with hand-worked graphs (two cliques joined by a bridge, a star, a chain cut into
pieces, isolated vertices dropped).
"""
from __future__ import annotations

from collections import Counter

import numpy as np

INF = 0xFFFFFFFF


def lpa(off, col, max_node=1024, min_node=2, rounds=10, max_degree=4096):
    """Synchronous label propagation on the undirected view (an edge counts once for
    each endpoint); a tie keeps the current label when it is among the most frequent,
    else takes the smallest; vertices with more than max_degree neighbours keep their
    label.  Then each label class, in id order, is cut into pieces of at most max_node
    members, pieces below min_node dropped, clusters numbered by smallest member."""
    n = len(off) - 1
    nbr = [[] for _ in range(n)]
    for u in range(n):
        for e in range(int(off[u]), int(off[u + 1])):
            v = int(col[e])
            nbr[u].append(v)
            nbr[v].append(u)
    label = list(range(n))
    for _ in range(rounds):
        nxt = list(label)
        for v in range(n):
            if not nbr[v] or len(nbr[v]) > max_degree:
                continue
            cnt = Counter(label[w] for w in nbr[v])
            best = max(cnt.values())
            if cnt.get(label[v], 0) == best:
                continue
            nxt[v] = min(lab for lab, c in cnt.items() if c == best)
        label = nxt
    classes = {}
    for v in range(n):
        classes.setdefault(label[v], []).append(v)
    pieces = []
    for lab in sorted(classes):
        mem = classes[lab]
        for s in range(0, len(mem), max_node):
            p = mem[s:s + max_node]
            if len(p) >= max(min_node, 1):
                pieces.append(p)
    pieces.sort(key=lambda p: p[0])
    out = np.full(n, INF, np.uint32)
    for i, p in enumerate(pieces):
        out[p] = i
    return out, len(pieces)
