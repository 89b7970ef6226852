"""The community clustering (suminc_cluster_lpa, host only) against the oracle's
restatement, and the oracle pinned by hand-worked graphs."""
import numpy as np
import pytest

import oracle
from oracle import clusters as OC
from suminc_amd import textio

INF = 0xFFFFFFFF


def _csr(n, edges):
    edges = sorted(set(edges))
    off = np.zeros(n + 1, np.uint64)
    for u, _ in edges:
        off[u + 1] += 1
    return np.cumsum(off).astype(np.uint64), np.array([v for _, v in edges], np.uint32)


def test_two_cliques_and_a_bridge():
    # cliques {0..4} and {5..9} (both directions), bridge 4 -> 5, isolated 10
    e = [(a, b) for a in range(5) for b in range(5) if a != b]
    e += [(a, b) for a in range(5, 10) for b in range(5, 10) if a != b]
    e += [(4, 5)]
    off, col = _csr(11, e)
    got, k = OC.lpa(off, col, 1024, 2, 10)
    assert k == 2
    assert list(got[:5]) == [0] * 5 and list(got[5:10]) == [1] * 5 and got[10] == INF


def test_star_and_a_chain_cut_into_pieces():
    """One round by hand.  Star 0 -> 1..6: every leaf takes 0; the centre sees six
    labels once each, not its own, and takes the smallest, 1.  Chain 7 -> ... -> 16:
    the ends take their only neighbour's label (7 -> 8, 16 -> 15), every inner vertex
    a tie of its two neighbours and the smaller (8 -> 7, 9 -> 8, ..., 15 -> 14).
    Classes: label 0 = {1..6}, cut by max_node 4 into [1..4] and [5, 6]; label 8 =
    {7, 9}; the rest single and dropped by min_node 2.  Numbered by first member."""
    e = [(0, v) for v in range(1, 7)] + [(v, v + 1) for v in range(7, 16)]
    off, col = _csr(17, e)
    got, k = OC.lpa(off, col, 4, 2, 1)
    want = [INF, 0, 0, 0, 0, 1, 1, 2, INF, 2] + [INF] * 7
    assert k == 3 and list(got) == want
    lib, kl = textio.cluster_lpa(off, col, 4, 2, 1)
    assert kl == 3 and list(lib) == want


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("max_node,min_node,rounds,max_degree", [(1024, 2, 10, 4096), (16, 3, 5, 64), (7, 2, 3, 8)])
def test_library_equals_oracle(seed, max_node, min_node, rounds, max_degree):
    g = oracle.Graph.rmat(10, 8, seed=seed)
    off, col, _ = g.csr()
    got, k = textio.cluster_lpa(off, col, max_node, min_node, rounds, max_degree)
    exp, ke = OC.lpa(off, col, max_node, min_node, rounds, max_degree)
    assert k == ke and np.array_equal(got, exp)


def test_round_trip_through_the_cluster_file(tmp_path):
    g = oracle.Graph.rmat(10, 8, seed=4)
    off, col, _ = g.csr()
    got, k = textio.cluster_lpa(off, col, 64, 2, 10)
    assert k > 1
    p = tmp_path / "g.e.c_64"
    textio.write_clusters(p, got)
    back, kb = textio.read_clusters(p, g.n, 2)
    assert kb == k and np.array_equal(back, got)
