"""The compressor with mirror vertices and the per-phase sketch entry points on the
device, against the oracle: the structure (node types 0..6, mirror map, cluster
lists, is_e_ / ib_e_ / sync_e_) bit-exact, the shortcut rows against the dense
closed form, one sketch round at a time in FP64 against the oracle's round, and
the whole query (sketch rounds, correction, plain rounds) at the plain fixed point."""
import numpy as np
import pytest

import oracle
from oracle import mirror as M
from suminc_amd import _lib as L
from tests.conftest import DAMPING
from tests.test_oracle_mirror import hand_graph, planted

pytestmark = pytest.mark.gpu
INF = 0xFFFFFFFF
FP32_ROWS = 2.0 ** -24 / (1 - DAMPING) * 8  # error budget of FP32 row weights


def _rmat_blocks():
    g = oracle.Graph.rmat(12, 16, seed=1)
    cl = np.where(np.arange(g.n) < 2048, np.arange(g.n) // 16, INF).astype(np.uint32)
    return g, cl


GRAPHS = {"hand": lambda: hand_graph(), "planted1": lambda: planted(1), "planted2": lambda: planted(2),
          "rmat12": _rmat_blocks}


def _rows(off, col, w=None):
    """CSR -> list of rows, sorted by column (order within a row carries no meaning)."""
    out = []
    for u in range(len(off) - 1):
        a, b = int(off[u]), int(off[u + 1])
        if w is None:
            out.append(sorted(int(x) for x in col[a:b]))
        else:
            out.append(sorted((int(x), float(y)) for x, y in zip(col[a:b], w[a:b])))
    return out


def _dense(rows, u, size):
    x = np.zeros(size)
    for c, w in rows[u]:
        x[c] += w
    return x


def _expected(g, cl, thr, k):
    off, col, _ = g.csr()
    st = M.compress(off, col, cl, 2, 1024, thr, k)
    return st, M.sketch_csr(st, M.shortcut_rows(st, DAMPING))


@pytest.mark.parametrize("name,thr,k", [("hand", 1.0, 1.0), ("planted1", 1.0, 4.0), ("planted1", 1.0, 1.0),
                                        ("planted2", 1.0, 4.0), ("planted2", 1.0, 1e8), ("rmat12", -1e9, 4.0),
                                        ("rmat12", -1e9, 1.0), ("rmat12", 1.0, 4.0)])
def test_compress_equals_oracle(gpu, name, thr, k):
    g, cl = GRAPHS[name]()
    off, col, _ = g.csr()
    st, csr = _expected(g, cl, thr, k)
    with gpu.Context(termcheck_threshold=1e-13, shortcut_max_rounds=20000) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        info = ctx.compress_mirror(cl, 2, 1024, thr, k)
        sk = ctx.sketch_download()
    n, a = g.n, st["all_node_num"]
    assert info["all_node_num"] == a and info["clusters"] == st["K"]
    assert info["in_mirrors"] == sum(map(len, st["in_ids"])) and info["out_mirrors"] == sum(map(len, st["out_ids"]))
    assert info["supernodes"] == len(st["sources"])
    assert info["node_type_count"] == np.bincount(st["node_type"], minlength=7).tolist()
    assert np.array_equal(sk["node_type"], np.array(st["node_type"], np.uint8))
    assert np.array_equal(sk["mirror_master"], np.array(st["master"], np.uint32))
    assert np.array_equal(sk["cluster_of"], np.array(st["cluster_of"], np.uint32))
    assert _rows(sk["ib_off"], sk["ib_col"]) == [sorted(r) for r in csr["ib"]]
    assert _rows(sk["sync_off"], sk["sync_col"]) == [sorted(r) for r in csr["sync"]]
    assert _rows(sk["sub_off"], sk["sub_col"]) == [sorted(r) for r in st["sub"]]
    # shortcut rows: the same rows, weights as the closed form's (FP32 at the
    # download; the device iterates to T/N per member, T = 1e-13)
    dis = _rows(sk["is_off"], sk["is_col"], sk["is_w"])
    div = _rows(sk["iv_off"], sk["iv_col"], sk["iv_w"])
    did = _rows(sk["id_off"], sk["id_col"], sk["id_w"])
    for u in range(a):
        assert ([c for c, _ in dis[u]] == [] ) == (csr["is"][u] == []), u
        e = _dense([sorted(r) for r in csr["is"]], u, a)
        assert np.allclose(_dense(dis, u, a), e, rtol=2e-6, atol=1e-9), u
        e = _dense([sorted(r) for r in csr["iv"]], u, n)
        assert np.allclose(_dense(div, u, n) + _dense(did, u, n), e, rtol=2e-6, atol=1e-9), u
        assert sum(abs(w) for _, w in did[u]) < 1e-10
    if name.startswith("planted") and k != 1e8:
        assert info["node_type_count"][5] > 0 and info["node_type_count"][6] > 0 and info["absorbed"] > 0


def _init_state(n, all_n):
    return (np.full(n, (1 - DAMPING) / n), np.zeros(n), np.zeros(all_n), np.zeros(all_n))


@pytest.mark.parametrize("name,k", [("hand", 1.0), ("planted1", 4.0), ("planted2", 1.0)])
def test_sketch_rounds_equal_oracle_round_by_round(gpu, name, k):
    """The oracle's sketch uploaded through suminc_upload_sketch (FP32 weights),
    then the precompute and rounds in FP64 on both sides: rank, delta and the
    residual agree after the precompute, every round and the correction."""
    g, cl = GRAPHS[name]()
    off, col, _ = g.csr()
    st, csr = _expected(g, cl, 1.0, k)
    sk = M.to_arrays(st, csr)
    T = 1e-9
    n, a = g.n, st["all_node_num"]
    delta, rank, bound, spnode = _init_state(n, a)
    eps = T / n
    o64 = off.astype(np.int64)
    with gpu.Context(pr_precision=gpu.PR_FP64, termcheck_threshold=T) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        ctx.upload_sketch(sk)
        ctx.pr_precompute()
        M.precompute(sk, o64, delta, rank, bound, spnode, DAMPING, T)
        r, d = ctx.pr_download()
        assert np.allclose(r, rank, rtol=1e-12, atol=1e-18)
        assert np.allclose(d, delta, rtol=1e-12, atol=1e-18)
        for rnd in range(12):
            res = ctx.pr_step_phase(L.PHASE_SKETCH)
            want = np.abs(delta).sum() + np.abs(bound).sum()
            assert res == pytest.approx(want, rel=1e-12), rnd
            M.sketch_round(sk, o64, delta, rank, bound, spnode, DAMPING, eps)
            r, d = ctx.pr_download()
            assert np.allclose(r, rank, rtol=1e-12, atol=1e-18), rnd
            assert np.allclose(d, delta, rtol=1e-12, atol=1e-18), rnd
        ctx.pr_correct()
        M.correct(sk, o64, delta, rank, bound, spnode, DAMPING, T)
        r, d = ctx.pr_download()
        assert np.allclose(r, rank, rtol=1e-11, atol=1e-17)
        assert np.allclose(d, delta, rtol=1e-11, atol=1e-17)
        # plain rounds from here reach the plain fixed point
        while ctx.pr_step_phase(L.PHASE_PLAIN) >= T:
            pass
        r, _ = ctx.pr_download()
    rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
    # the rows went through FP32 weights: each use moves its mass by <= 2^-24
    # relative, at most the total mass 1 over (1 - d) rounds' worth
    assert np.abs(r - rs).sum() < T / (1 - DAMPING) + FP32_ROWS


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name,thr,k", [("planted1", 1.0, 4.0), ("planted2", 1.0, 1.0), ("rmat12", -1e9, 4.0),
                                        ("rmat12", 1.0, 1e8)])
def test_sketch_query_equals_fixed_point(gpu, name, thr, k, prec):
    g, cl = GRAPHS[name]()
    off, col, _ = g.csr()
    rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
    T = 1e-10 * g.n
    p = gpu.PR_FP64 if prec == "fp64" else gpu.PR_FP32_KAHAN
    with gpu.Context(pr_precision=p, termcheck_threshold=T) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        info = ctx.compress_mirror(cl, 2, 1024, thr, k)
        st = ctx.sketch_pr_run()
        r, _ = ctx.pr_download()
    err = np.abs(r - rs).sum()
    # the correction drops what the exits still hold below T / 1000
    assert err <= (st["residual"] + 1e-3 * T) / (1 - DAMPING) + (1e-11 if prec == "fp64" else 2e-6), (err, st)
    if info["clusters"]:
        assert 0 < st["sketch_rounds"] < 1000 and st["precompute_edges"] > 0, st
    assert st["rounds"] == st["sketch_rounds"] + st["residual_rounds"]


def test_phase_by_phase_equals_the_query(gpu):
    """The worker's phase order driven through the per-phase entry points gives
    what suminc_sketch_pr_run gives."""
    g, cl = planted(3)
    off, col, _ = g.csr()
    T = 1e-10 * g.n
    with gpu.Context(pr_precision=gpu.PR_FP64, termcheck_threshold=T) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        ctx.compress_mirror(cl, 2, 1024, 1.0, 4.0)
        st = ctx.sketch_pr_run()
        r1, _ = ctx.pr_download()
        ctx.pr_init()
        ctx.pr_precompute()
        rounds = 0
        while ctx.pr_step_phase(L.PHASE_SKETCH) >= T:
            rounds += 1
        ctx.pr_correct()
        plain = 0
        while ctx.pr_step_phase(L.PHASE_PLAIN) >= T:
            plain += 1
        r2, _ = ctx.pr_download()
    # the query checks the residual before a round; the loop above ran one round past it
    assert rounds == st["sketch_rounds"]
    assert abs(plain - st["residual_rounds"]) <= 1
    assert np.abs(r1 - r2).sum() < 2 * T / (1 - DAMPING)


def test_download_upload_round_trip(gpu):
    g, cl = planted(1)
    off, col, _ = g.csr()
    T = 1e-10 * g.n
    rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
    with gpu.Context(pr_precision=gpu.PR_FP64, termcheck_threshold=T) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        ctx.compress_mirror(cl, 2, 1024, 1.0, 4.0)
        sk = ctx.sketch_download()
    with gpu.Context(pr_precision=gpu.PR_FP64, termcheck_threshold=T) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        ctx.upload_sketch(sk)
        st = ctx.sketch_pr_run()
        r, _ = ctx.pr_download()
    assert st["sketch_rounds"] > 0
    assert np.abs(r - rs).sum() <= st["residual"] / (1 - DAMPING) + FP32_ROWS  # FP32 weights in the download


def test_upload_sketch_rejects_bad_input(gpu):
    g, cl = planted(1)
    off, col, _ = g.csr()
    st, csr = _expected(g, cl, 1.0, 4.0)
    good = M.to_arrays(st, csr)
    n = g.n
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        with pytest.raises(gpu.SumincError) as e:
            ctx.pr_step_phase(L.PHASE_SKETCH)
        assert e.value.code == gpu.E_STATE  # nothing uploaded yet

        def bad(**change):
            sk = dict(good)
            sk.update(change)
            with pytest.raises(gpu.SumincError) as e:
                ctx.upload_sketch(sk)
            assert e.value.code == gpu.E_INVALID, change.keys()

        t = good["node_type"].copy()
        t[0] = 7
        bad(node_type=t)
        m = good["mirror_master"].copy()
        m[0] = n
        bad(mirror_master=m)
        c = good["is_col"].copy()
        c[0] = st["all_node_num"]
        bad(is_col=c)
        o = good["ib_off"].copy()
        o[5] = o[6] + 1  # decreasing row pointers
        bad(ib_off=o)
        # a subgraph row on a Single
        v = int(np.nonzero(good["node_type"] == 0)[0][0])
        so = good["sub_off"].astype(np.int64)
        cols = list(good["sub_col"])
        cols.insert(int(so[v]), 0)
        so[v + 1:] += 1
        bad(sub_off=so.astype(np.uint64), sub_col=np.array(cols, np.uint32))
        # a sync row on a vertex that is not a master
        v = int(np.nonzero(good["node_type"] == 4)[0][0])
        so = good["sync_off"].astype(np.int64)
        cols = list(good["sync_col"])
        cols.insert(int(so[v]), n)
        so[v + 1:] += 1
        bad(sync_off=so.astype(np.uint64), sync_col=np.array(cols, np.uint32))
        with pytest.raises(gpu.SumincError) as e:
            ctx.pr_step_phase(7)
        assert e.value.code == gpu.E_INVALID
        # the good one uploads; after a batch the graph has moved on
        ctx.upload_sketch(good)
        ctx.pr_step_phase(L.PHASE_SKETCH)
        ops = g.gen_batch(seed=1, batch_id=0, n_del=5, n_ins=5)
        ctx.apply_batch(ops)
        with pytest.raises(gpu.SumincError) as e:
            ctx.pr_step_phase(L.PHASE_SKETCH)
        assert e.value.code == gpu.E_STATE


@pytest.mark.parametrize("weights", [None, "u32"])
def test_sssp_step_equals_dijkstra(gpu, weights):
    g = oracle.Graph.rmat(13, 16, seed=2, weights=weights, wmin=1, wmax=64)
    off, col, w = g.csr()
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col, w)
        ctx.sssp_init(0)
        steps = 0
        while ctx.sssp_step():
            steps += 1
            assert steps < 100000
        dist, parent = ctx.sssp_download()
    exp, _ = g.dijkstra(0)
    assert np.array_equal(dist, exp)
    assert steps > 3


@pytest.mark.parametrize("name,k", [("hand", 1.0), ("planted1", 4.0), ("planted2", 1.0)])
def test_cluster_cache_equals_oracle_and_reloads(gpu, tmp_path, name, k):
    """suminc_compress_mirror_save writes the oracle's bytes (the reference's layout);
    suminc_compress_mirror_load on a fresh context gives the same products."""
    g, cl = GRAPHS[name]()
    off, col, _ = g.csr()
    st = M.compress(off, col, cl, 2, 1024, 1.0, k)
    path = tmp_path / "clusters.bin"
    with gpu.Context(termcheck_threshold=1e-13, shortcut_max_rounds=20000) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        info = ctx.compress_mirror(cl, 2, 1024, 1.0, k)
        sk = ctx.sketch_download()
        ctx.compress_mirror_save(path)
    assert path.read_bytes() == M.cluster_cache_bytes(st)
    with gpu.Context(termcheck_threshold=1e-13, shortcut_max_rounds=20000) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        info2 = ctx.compress_mirror_load(path)
        sk2 = ctx.sketch_download()
        # a damaged cache is refused
        raw = path.read_bytes()
        bad = tmp_path / "bad.bin"
        for blob in (raw[:-4], raw[:8] + (int.from_bytes(raw[8:12], "little") + 1).to_bytes(4, "little") + raw[12:],
                     raw[:-8 - 4 * g.n] + (5).to_bytes(4, "little") + raw[-4 - 4 * g.n:]):
            bad.write_bytes(blob)
            with pytest.raises(gpu.SumincError) as e:
                ctx.compress_mirror_load(bad)
            assert e.value.code == gpu.E_INVALID
    assert {k2: v for k2, v in info2.items() if k2 != "shortcut_work"} == \
        {k2: v for k2, v in info.items() if k2 != "shortcut_work"}
    for key, v in sk.items():
        if key.endswith("_w"):
            assert np.allclose(sk2[key], v, rtol=1e-6, atol=1e-12), key
        else:
            assert np.array_equal(np.asarray(sk2[key]), np.asarray(v)), key


def test_community_clusters_through_the_query(gpu):
    """Label-propagation clusters (suminc_cluster_lpa) into the compressor with
    mirrors and the whole query: the plain fixed point."""
    from suminc_amd import textio

    g, _ = planted(1)
    off, col, _ = g.csr()
    cl, k = textio.cluster_lpa(off, col, 1024, 2, 10)
    assert k >= 8
    T = 1e-10 * g.n
    rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
    with gpu.Context(pr_precision=gpu.PR_FP64, termcheck_threshold=T) as ctx:
        ctx.upload_csr(off, col)
        ctx.pr_init()
        info = ctx.compress_mirror(cl, 2, 1024, 1.0, 4.0)
        st = ctx.sketch_pr_run()
        r, _ = ctx.pr_download()
    assert info["clusters"] > 0
    assert np.abs(r - rs).sum() <= (st["residual"] + 1e-3 * T) / (1 - DAMPING) + 1e-11
