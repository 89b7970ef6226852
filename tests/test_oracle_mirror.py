"""The oracle's restatement of the compressor with mirror vertices and of the sketch
iteration, pinned without a GPU:

* a hand-worked example whose every product (mirror selection by mirror_k, the
  four benefits, absorption of unclustered candidates, mirror ids, subgraph rows
  with entry priority, exits, node types 5 and 6, is_e_ / ib_e_ / sync_e_) is
  written out below from the rules of compressor_base.h:3253-4899;
* mirror_k = 1e8 reduces to the mirror-free compressor, whose C restatement is
  pinned by its own tests;
* the shortcut rows of the closed form equal the C oracle's iterated shortcut
  where there are no mirrors;
* the sketch iteration (rounds by node type, out-mirror sync, correction, then
  plain rounds) reaches the plain system's fixed point: a wrong row, sign or
  order of passes moves it."""
import numpy as np
import pytest

import oracle
from oracle import mirror as M
from tests.conftest import DAMPING

INF = 0xFFFFFFFF

# Hand-worked graph: cluster A = {0,1,2,3}, cluster B = {4,5,6}, 7, 8, 9 unclustered.
HAND_EDGES = [
    (0, 1), (1, 2), (2, 3), (3, 0), (1, 3),          # A internal (5)
    (4, 5), (5, 6), (6, 4), (4, 6), (5, 4), (6, 5),  # B internal (6)
    (7, 0), (7, 1), (7, 2),                          # 7 -> A three times: in-mirror candidate, absorbed
    (4, 0), (4, 2),                                  # 4 (in B) -> A twice: in-mirror, stays (master)
    (6, 0), (6, 1),                                  # 6 (in B) -> A twice: in-mirror, stays (master)
    (2, 8), (3, 8),                                  # A -> 8 twice: out-mirror candidate, absorbed
    (0, 5), (1, 5),                                  # A -> 5 (in B) twice: out-mirror, stays
    (3, 9),                                          # A -> 9 once: no mirror, 3 is an exit
    (9, 4),                                          # 9 -> B: 4 is an entry of B
]


def hand_graph():
    src, dst = zip(*HAND_EDGES)
    g = oracle.Graph.from_edges(10, src, dst)
    cl = np.array([0, 0, 0, 0, 1, 1, 1, INF, INF, INF], np.uint32)
    return g, cl


def test_hand_example_structure():
    """Every product worked out by hand.

    A: inner 5; in_freq {7:3, 4:2, 6:2} > 1 -> in-mirrors {4,6,7} (7 in-edges);
    out_freq {5:2, 8:2, 9:1} -> out-mirrors {5,8} (4 out-edges); old exits
    {0,1,2,3}, old entries {0,1,2}; new exits {3}, no new entries.  Benefits:
    5-12 = -7, 12-(4*3+3) = -3, 9-(3*3+2) = -2, 16-(3*3+5) = 2 -> both kinds,
    accepted (2 > 1).  B (after A): the edges 0->5, 1->5 are carried by A's
    out-mirror of 5 and 4->A, 6->A by A's in-mirrors, so no frequency counts them;
    inner 6, old entries {4,5}, old exits {4,6}: all four benefits 6-4 = 2 > 1,
    the first wins -> no mirrors, entries {4,5}.  Final build of A: 7 and 8 are
    unclustered and absorbed (neither has an in-edge from outside A), 4 and 6
    stay in-mirrors (ids 10, 11), 5 an out-mirror (id 12)."""
    g, cl = hand_graph()
    off, col, _ = g.csr()
    st = M.compress(off, col, cl, 2, 1024, 1.0, 1.0)
    assert st["K"] == 2 and st["all_node_num"] == 13
    assert st["members"] == [[0, 1, 2, 3, 7, 8], [4, 5, 6]]
    assert st["ids"] == [[0, 1, 2, 3, 7, 8, 10, 11, 12], [4, 5, 6]]
    assert st["master"] == [4, 6, 5]
    assert st["in_ids"] == [[10, 11], []] and st["out_ids"] == [[12], []]
    assert st["sources"] == [10, 11, 4, 5] and st["source_cluster"] == [0, 0, 1, 1]
    assert st["cluster_of"] == [0, 0, 0, 0, 1, 1, 1, 0, 0, INF, 0, 0, 0]
    # node types: 0..2, 7, 8 Inner; 3 OnlyOut; 4 entry+exit+master -> BothOutInMaster;
    # 5 OnlyIn; 6 exit+master -> OutMaster; 9 Single
    assert st["node_type"] == [4, 4, 4, 1, 6, 2, 5, 4, 4, 0]
    sub = st["sub"]
    assert sub[0] == [1, 12] and sub[1] == [2, 3, 12]  # the out-mirror of 5 takes 0->5, 1->5
    assert sub[2] == [3, 8] and sub[3] == [0, 8] and sub[7] == [0, 1, 2] and sub[8] == []
    assert sub[10] == [0, 2] and sub[11] == [0, 1]     # in-mirrors: their masters' edges into A
    assert sorted(sub[4]) == [5, 6] and sorted(sub[5]) == [4, 6] and sorted(sub[6]) == [4, 5]
    assert sub[12] == [] and sub[9] == []
    rows = M.shortcut_rows(st, DAMPING)
    csr = M.sketch_csr(st, rows)
    # ib_e_: 9 (Single) all out-edges; 3 only ->9 (->0, ->8 stay in A); 4 nothing
    # (->0, ->2 are carried by its in-mirror in A); 6 nothing (likewise)
    assert csr["ib"][9] == [4] and csr["ib"][3] == [9] and csr["ib"][4] == [] and csr["ib"][6] == []
    assert all(csr["ib"][v] == [] for v in (0, 1, 2, 5, 7, 8))
    # sync_e_: masters -> in-mirrors, out-mirror -> master
    assert csr["sync"][4] == [10] and csr["sync"][6] == [11] and csr["sync"][12] == [5]
    assert sum(len(r) for r in csr["sync"]) == 3
    # is_e_ rows: in-mirror sources and entries (4: BothOutInMaster, 5: OnlyIn)
    assert [u for u in range(13) if csr["is"][u]] == [4, 5, 10, 11]
    assert [x for x, _ in csr["is"][10]] == [3, 12]  # A's exit 3 and its out-mirror
    assert [x for x, _ in csr["is"][4]] == [4, 6] and [x for x, _ in csr["is"][5]] == [4, 6]


def test_hand_example_shortcut_values():
    """In-mirror 11 (master 6, outdeg 4) sends d/4 to 0 and to 1; from there the
    rows solve r = x0 + A r over A's members: check two entries by hand."""
    g, cl = hand_graph()
    off, col, _ = g.csr()
    st = M.compress(off, col, cl, 2, 1024, 1.0, 1.0)
    rows = M.shortcut_rows(st, DAMPING)
    d = DAMPING
    # closed form of A's internal chain: r1 = x1 + d/3 r0 ... verified as a system
    mem = st["members"][0]
    deg = np.diff(off)
    A = np.zeros((6, 6))
    for a, v in enumerate(mem):
        for x in st["sub"][v]:
            if x in mem:
                A[mem.index(x), a] += d / deg[v]
    x0 = np.zeros(6)
    x0[0] += d / 4
    x0[1] += d / 4
    r = np.linalg.solve(np.eye(6) - A, x0)
    iv = dict(rows[11]["inner_value"])
    bd = dict(rows[11]["bound_delta"])
    assert iv[0] == pytest.approx(r[0], rel=1e-14) and bd[3] == pytest.approx(r[3], rel=1e-14)
    # the out-mirror of 5 collects what 0 (out-degree 2: 1, 5) and 1 (out-degree 3:
    # 2, 3, 5) send along their edges to 5: d/2 of 0's reach and d/3 of 1's
    assert deg[0] == 2 and deg[1] == 3
    assert bd[12] == pytest.approx(d / 2 * r[0] + d / 3 * r[1], rel=1e-14)
    assert bd[12] != pytest.approx(d / 3 * r[0] + d / 3 * r[1], rel=1e-6)


def planted(seed, n_cl=16, size=32, n_free=64, fan=8, links=3, noise=0.1):
    """Planted clusters that want mirrors: n_cl clusters of `size` vertices, each
    member with `fan` out-edges inside; per cluster `links` pairs of a member with
    `fan` edges into another cluster (an in-mirror there) and `fan` members of that
    cluster with edges to one member (an out-mirror there); n_free unclustered
    vertices, half of them with `fan` edges into or out of one cluster (absorbed
    candidates); a `noise` share of members with one random out-edge."""
    rng = np.random.default_rng(seed)
    n = n_cl * size + n_free
    e = []
    for c in range(n_cl):
        for i in range(size):
            e += [(c * size + i, c * size + int(v)) for v in rng.choice(size, fan, replace=False)]
    for c in range(n_cl):
        for _ in range(links):
            y = int(rng.integers(n_cl))
            if y == c:
                continue
            w = c * size + int(rng.integers(size))
            e += [(w, y * size + int(v)) for v in rng.choice(size, fan, replace=False)]
            t = c * size + int(rng.integers(size))
            e += [(y * size + int(v), t) for v in rng.choice(size, fan, replace=False)]
    for f in range(n_free):
        u = n_cl * size + f
        y = int(rng.integers(n_cl))
        for v in rng.choice(size, fan if f % 2 == 0 else 2, replace=False):
            e.append((u, y * size + int(v)) if f % 4 < 2 else (y * size + int(v), u))
        e += [(u, int(rng.integers(n))), (int(rng.integers(n)), u)]
    for u in range(n_cl * size):
        if rng.random() < noise:
            e.append((u, int(rng.integers(n))))
    e = np.unique(np.array([x for x in e if x[0] != x[1]], np.int64), axis=0)
    g = oracle.Graph.from_edges(n, e[:, 0], e[:, 1])
    cl = np.where(np.arange(n) < n_cl * size, np.arange(n) // size, INF).astype(np.uint32)
    return g, cl


def test_planted_graph_exercises_every_rule():
    g, cl = planted(1)
    off, col, _ = g.csr()
    st = M.compress(off, col, cl, 2, 1024, 1.0, 4.0)
    types = np.bincount(st["node_type"], minlength=7)
    assert np.all(types[1:] > 0), types  # every node type but Single (absorbed free vertices: some Singles)
    assert len(st["in_ids"]) and sum(map(len, st["in_ids"])) > 20 and sum(map(len, st["out_ids"])) > 20
    assert sum(len(m) for m in st["members"]) > 16 * 32  # unclustered candidates absorbed
    # without mirrors almost nothing is accepted on this graph
    assert M.compress(off, col, cl, 2, 1024, 1.0, 1e8)["K"] < 3


def _blocks(n, upto, size):
    return np.where(np.arange(n) < upto, np.arange(n) // size, INF).astype(np.uint32)


@pytest.mark.parametrize("thr", [1.0, -1e9])
def test_mirror_free_equals_c_compressor(thr):
    g = oracle.Graph.rmat(11, 8, seed=4)
    off, col, _ = g.csr()
    raw = _blocks(g.n, 1024, 16)
    st = M.compress(off, col, raw, 2, 1024, thr, 1e8)
    cid, typ = g.sum_structure(raw, 2, 1024, thr)
    assert st["all_node_num"] == g.n and not st["master"]
    assert np.array_equal(np.array(st["node_type"], np.uint8), typ)
    mine = np.array(st["cluster_of"][:g.n], np.uint64)
    assert np.array_equal(mine == INF, cid == INF)
    # the same partition (ids may be numbered differently)
    pairs = {(int(a), int(b)) for a, b in zip(mine, cid) if a != INF}
    assert len(pairs) == len({a for a, _ in pairs}) == len({b for _, b in pairs})


def test_mirror_free_rows_equal_c_shortcut():
    g = oracle.Graph.rmat(10, 8, seed=4)
    off, col, _ = g.csr()
    raw = _blocks(g.n, 512, 16)
    st = M.compress(off, col, raw, 2, 1024, -1e9, 1e8)
    cid, _ = g.sum_structure(raw, 2, 1024, -1e9)
    rows = M.shortcut_rows(st, DAMPING)
    checked = 0
    for s in st["sources"][:30]:
        r, dl, _ = g.sum_shortcut(cid, s, DAMPING, 1e-300, 100000)
        got = np.zeros(g.n)
        for x, w in rows[s]["bound_delta"] + rows[s]["inner_value"]:
            got[x] = w
        assert np.abs(got - (r + dl)).max() < 1e-12, s
        checked += 1
    assert checked == 30


def _run_sketch(g, sk, T, max_rounds=5000):
    off, _, _ = g.csr()
    n, all_n = int(sk["n"]), int(sk["all_node_num"])
    delta = np.full(n, (1 - DAMPING) / n)
    rank = np.zeros(n)
    bound = np.zeros(all_n)
    spnode = np.zeros(all_n)
    M.precompute(sk, off.astype(np.int64), delta, rank, bound, spnode, DAMPING, T)
    rounds = 0
    while np.abs(delta).sum() + np.abs(bound).sum() >= T and rounds < max_rounds:
        M.sketch_round(sk, off.astype(np.int64), delta, rank, bound, spnode, DAMPING, T / n)
        rounds += 1
    M.correct(sk, off.astype(np.int64), delta, rank, bound, spnode, DAMPING, T)
    return rank, delta, rounds


@pytest.mark.parametrize("mirror_k", [1.0, 4.0])
def test_sketch_iteration_reaches_fixed_point(mirror_k):
    """Sketch rounds + correction + plain rounds = the plain fixed point (rows in
    FP64, so the only error left is the stopping rule's)."""
    g, cl = planted(2)
    off, col, _ = g.csr()
    st = M.compress(off, col, cl, 2, 1024, 1.0, mirror_k)
    assert len(st["master"]) > 40
    assert any(t in (5, 6) for t in st["node_type"])
    csr = M.sketch_csr(st, M.shortcut_rows(st, DAMPING))
    sk = M.to_arrays(st, csr)
    for k in ("is_w", "iv_w", "id_w"):  # keep FP64 weights for the check
        sk[k] = np.array([w for row in csr[k[:2]] for _, w in row], np.float64)
    T = 1e-9
    rank, delta, rounds = _run_sketch(g, sk, T)
    assert 3 < rounds < 5000  # converged by the residual (the precompute moved the members' mass)
    r, dl, _ = g.pr_solve(DAMPING, tol=1e-15, r=rank, delta=delta)
    rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
    assert np.abs(r - rs).sum() < 1e-12


def test_sketch_iteration_catches_a_wrong_row():
    """Planting a mistake in the oracle's rows (dropping one is_e_ entry) moves the
    fixed point: the check above would see it."""
    g, cl = planted(2)
    off, col, _ = g.csr()
    st = M.compress(off, col, cl, 2, 1024, 1.0, 4.0)
    csr = M.sketch_csr(st, M.shortcut_rows(st, DAMPING))
    u = next(u for u in range(st["all_node_num"]) if len(csr["is"][u]) > 1)
    csr["is"][u] = csr["is"][u][1:]
    sk = M.to_arrays(st, csr)
    for k in ("is_w", "iv_w", "id_w"):
        sk[k] = np.array([w for row in csr[k[:2]] for _, w in row], np.float64)
    rank, delta, _ = _run_sketch(g, sk, 1e-9)
    r, _, _ = g.pr_solve(DAMPING, tol=1e-15, r=rank, delta=delta)
    rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
    assert np.abs(r - rs).sum() > 1e-9


def test_hand_example_cluster_cache_bytes():
    """The cluster cache of the hand example, word by word: 5 mirrors (A: in 4, 6, 7;
    out 5, 8), A = {0..3} with no sources, B = {4, 5, 6} with sources 4, 5; Fc and
    id2spids for the ten vertices; the end marker."""
    import struct
    g, cl = hand_graph()
    off, col, _ = g.csr()
    st = M.compress(off, col, cl, 2, 1024, 1.0, 1.0)
    words = [5, 2,
             4, 0, 1, 2, 3, 4, 0, 1, 2, 3, 0, 3, 4, 6, 7, 2, 5, 8,
             3, 4, 5, 6, 3, 4, 5, 6, 2, 4, 5, 0, 0]
    m = 2**31 - 1
    fc = [-1, -1, -1, -1, -2, -2, -2, m, m, m]
    sp = [0, 0, 0, 0, 1, 1, 1, INF, INF, INF]
    want = struct.pack(f"<{len(words)}I", *words) + struct.pack("<10i", *fc) + struct.pack("<10I", *sp)
    want += struct.pack("<i", -1)
    assert M.cluster_cache_bytes(st) == want
