"""The reference's text formats (edge, update and cluster files) through libsuminc's
host parsers.  CPU only: these calls do no device work.

Expected values are built here from the format definitions with plain Python
(dict / list bookkeeping over the same lines), independent of the C++ parser:
  * edge files: e_fragment_loader.h:107-170 (skip empty and '#' lines, skip lines
    that do not parse, both directions when undirected); a repeated pair keeps its
    first line (the device graph holds one edge per pair);
  * update files: inc_fragment_builder.h:29-72 ("a u v [w]" / "d u v", other types
    rejected);
  * cluster files: compressor_base.h:3153-3250 (records "<size> <ids>", a std::set
    per record, records below min_node_num dropped, kept ones numbered in order).
"""
from pathlib import Path

import numpy as np
import pytest

from suminc_amd import _lib as L
from suminc_amd import textio

GOLD = Path(__file__).resolve().parent / "golden"


def csr_from_pairs(pairs, n):
    """CSR with sorted rows of a dict {(u, v): w} (insertion order irrelevant)."""
    rows = [[] for _ in range(n)]
    for (u, v), w in pairs.items():
        rows[u].append((v, w))
    off = np.zeros(n + 1, np.uint64)
    col, wt = [], []
    for u in range(n):
        rows[u].sort()
        off[u + 1] = off[u] + len(rows[u])
        for v, w in rows[u]:
            col.append(v)
            wt.append(w)
    return off, np.array(col, np.uint32), wt


def test_golden_edge_file_matches_loadtxt():
    with textio.EdgeFile(GOLD / "test_base.txt") as ef:
        off, col, w = ef.csr()
        assert w is None and ef.skipped_lines == 0
    e = np.loadtxt(GOLD / "test_base.txt", dtype=np.int64, comments="#")
    n = int(e.max()) + 1
    pairs = {}
    for u, v in e:
        pairs.setdefault((int(u), int(v)), 0)
    o2, c2, _ = csr_from_pairs(pairs, n)
    assert ef.n == n and ef.m == len(pairs)
    assert np.array_equal(off, o2) and np.array_equal(col, c2)


@pytest.mark.parametrize("directed", [True, False])
def test_edge_file_rules(tmp_path, directed):
    rng = np.random.default_rng(7)
    lines, pairs, skipped = [], {}, 0
    n_true = 0
    for i in range(4000):
        r = rng.random()
        if r < 0.03:
            lines.append("# a comment line")
            continue
        if r < 0.05:
            lines.append("")
            continue
        if r < 0.07:
            lines.append(f"{rng.integers(0, 500)} x{rng.integers(0, 9)} 3")  # does not parse
            skipped += 1
            continue
        if r < 0.08:
            lines.append(f"{rng.integers(0, 500)}")  # missing destination
            skipped += 1
            continue
        u, v = int(rng.integers(0, 500)), int(rng.integers(0, 500))
        if rng.random() < 0.1 and pairs:
            u, v = list(pairs)[int(rng.integers(0, len(pairs)))]  # a repeated pair
        w = int(rng.integers(1, 100))
        sep = "\t" if rng.random() < 0.3 else " "
        end = "\r" if rng.random() < 0.1 else ""
        lines.append(f"{u}{sep}{v}{sep}{w}{end}")
        pairs.setdefault((u, v), w)  # the first line of a pair wins
        if not directed:
            pairs.setdefault((v, u), w)
        n_true = max(n_true, u + 1, v + 1)
    path = tmp_path / "g.e"
    path.write_text("\n".join(lines) + "\n")
    with textio.EdgeFile(path, weights="u32", directed=directed) as ef:
        off, col, w = ef.csr()
        assert ef.skipped_lines == skipped
        assert ef.n == n_true
    o2, c2, w2 = csr_from_pairs(pairs, n_true)
    assert np.array_equal(off, o2) and np.array_equal(col, c2)
    assert np.array_equal(w, np.array(w2, np.uint32))


def test_edge_file_float_weights_and_no_trailing_newline(tmp_path):
    path = tmp_path / "f.e"
    path.write_text("0 1 0.5\n1 2 1e-3\n2 0 7\n0 1 9.0")  # repeated (0, 1): first kept
    with textio.EdgeFile(path, weights="f32") as ef:
        off, col, w = ef.csr()
    assert off.tolist() == [0, 1, 2, 3]
    assert col.tolist() == [1, 2, 0]
    assert w.tolist() == pytest.approx([0.5, 1e-3, 7.0])


def test_edge_file_zero_integer_weight_line_is_skipped(tmp_path):
    path = tmp_path / "z.e"
    path.write_text("0 1 0\n1 2 4\n")
    with textio.EdgeFile(path, weights="u32") as ef:
        assert ef.skipped_lines == 1 and ef.m == 1


def test_missing_file_reports_through_last_error(tmp_path):
    with pytest.raises(L.SumincError) as ei:
        textio.EdgeFile(tmp_path / "nope.e")
    assert ei.value.code == L.E_INVALID and "cannot open" in str(ei.value)


@pytest.mark.parametrize("directed", [True, False])
def test_update_file(tmp_path, directed):
    text = "# batch\na 0 5 3\nd 2 3\n\na 7 7 1\nd 4 1 99\n"
    path = tmp_path / "u.e"
    path.write_text(text)
    ops = textio.read_updates(path, weights="u32", directed=directed)
    want = [(L.OP_ADD, 0, 5, 3), (L.OP_DEL, 2, 3, 0), (L.OP_ADD, 7, 7, 1), (L.OP_DEL, 4, 1, 0)]
    if not directed:  # both directions, a self-loop once
        want = [want[0], (L.OP_ADD, 5, 0, 3), want[1], (L.OP_DEL, 3, 2, 0), want[2], want[3], (L.OP_DEL, 1, 4, 0)]
    got = [(int(o["type"]), int(o["u"]), int(o["v"]), int(o["w_u32"])) for o in ops]
    assert got == want


def test_update_file_errors_name_the_line(tmp_path):
    path = tmp_path / "bad.e"
    path.write_text("a 0 1 2\nx 1 2\n")
    with pytest.raises(L.SumincError) as ei:
        textio.read_updates(path, weights="u32")
    assert "line 2" in str(ei.value) and "invalid op" in str(ei.value)
    path.write_text("a 0 1\n")
    with pytest.raises(L.SumincError) as ei:
        textio.read_updates(path, weights="u32")
    assert "missing weight" in str(ei.value)
    assert len(textio.read_updates(path, weights=None)) == 1


def test_cluster_file_read(tmp_path):
    # records may span lines; duplicates inside a record count once
    path = tmp_path / "g.e.c_5"
    path.write_text("2 0 1\n3 3 4\n5\n1 9\n2 6 6\n4 2 7 8 10\n")
    cof, k = textio.read_clusters(path, 12, min_node_num=2)
    inf = L.INF_U32
    assert k == 3
    assert cof.tolist() == [0, 0, 2, 1, 1, 1, inf, 2, 2, inf, 2, inf]
    cof1, k1 = textio.read_clusters(path, 12, min_node_num=1)
    assert k1 == 5 and cof1[9] == 2 and cof1[6] == 3 and cof1[2] == 4


def test_cluster_file_errors(tmp_path):
    path = tmp_path / "bad.c_5"
    path.write_text("3 0 1\n")
    with pytest.raises(L.SumincError, match="ends before"):
        textio.read_clusters(path, 10)
    path.write_text("2 0 11\n")
    with pytest.raises(L.SumincError, match=">= n"):
        textio.read_clusters(path, 10)


def test_cluster_file_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    n = 3000
    cof = rng.integers(0, 200, n).astype(np.uint32)
    cof[rng.random(n) < 0.3] = L.INF_U32
    path = tmp_path / "rt.c_1024"
    textio.write_clusters(path, cof)
    back, k = textio.read_clusters(path, n, min_node_num=1)
    # the file numbers the non-empty clusters in id order
    used = np.unique(cof[cof != L.INF_U32])
    remap = {int(c): i for i, c in enumerate(used)}
    want = np.array([remap[int(c)] if c != L.INF_U32 else L.INF_U32 for c in cof], np.uint32)
    assert k == len(used) and np.array_equal(back, want)
    for line in path.read_text().splitlines():
        t = [int(x) for x in line.split()]
        assert t[0] == len(t) - 1 and t[1:] == sorted(t[1:])


def test_cluster_file_name_follows_the_compressor():
    assert textio.cluster_file_path("data/road_usa.e", 5000) == "data/road_usa.e.c_5000"
    assert textio.cluster_file_path("data/uk_w.e", 1000) == "data/uk.e.c_1000"
    assert textio.cluster_file_path("g_ud.random.e", 8) == "g.e.c_8"
