"""Device CSR/CSC construction, batch generation and batch application vs the
oracle: bit-exact (integer work)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def csc_from_csr(off, col, w):
    n = len(off) - 1
    src = np.repeat(np.arange(n), np.diff(off).astype(np.int64))
    order = np.lexsort((src, col))
    ioff = np.zeros(n + 1, np.int64)
    np.add.at(ioff, col.astype(np.int64) + 1, 1)
    return np.cumsum(ioff), src[order], (None if w is None else w[order])


@pytest.mark.parametrize("scale,ef,weights,seed", [(6, 4, None, 1), (12, 16, None, 1), (12, 16, "u32", 2),
                                                   (13, 16, "f32", 4), (16, 16, None, 1)])
def test_rmat_equals_oracle(gpu, scale, ef, weights, seed):
    with gpu.Context() as ctx:
        ctx.generate_rmat(scale, ef, seed=seed, weights=weights, wmin=1, wmax=64)
        off, col, w = ctx.download_csr()
    g = oracle.Graph.rmat(scale, ef, seed=seed, weights=weights, wmin=1, wmax=64)
    o_off, o_col, o_w = g.csr()
    assert np.array_equal(off, o_off)
    assert np.array_equal(col, o_col)
    if weights:
        assert np.array_equal(w.view(np.uint32), o_w.view(np.uint32))


def test_upload_roundtrip_and_errors(gpu):
    g = oracle.Graph.rmat(10, 8, seed=5, weights="u32")
    off, col, w = g.csr()
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col, w)
        o2, c2, w2 = ctx.download_csr()
        assert np.array_equal(o2, off) and np.array_equal(c2, col) and np.array_equal(w2, w)
        bad = col.copy()
        bad[3] = g.n + 7
        with pytest.raises(gpu.SumincError) as e:
            ctx.upload_csr(off, bad, w)
        assert e.value.code == -1
        with pytest.raises(gpu.SumincError):
            ctx.pr_run()  # state not initialised after a failed upload


def test_empty_and_edgeless_graphs(gpu):
    with gpu.Context() as ctx:
        ctx.upload_csr(np.zeros(5, np.uint64), np.zeros(0, np.uint32))
        ctx.pr_init()
        st = ctx.pr_run()
        r, d = ctx.pr_download()
        # isolated vertices keep their initial delta as rank after one round
        assert np.allclose(r, (1 - ctx.params.damping) / 4)
        ctx.sssp_init(2)
        ctx.sssp_run()
        dist, par = ctx.sssp_download()
        assert dist.tolist() == [0xFFFFFFFF, 0xFFFFFFFF, 0, 0xFFFFFFFF]
        assert par.tolist() == [0, 1, 2, 3]


@pytest.mark.parametrize("weights", [None, "u32"])
def test_generate_batch_equals_oracle(gpu, weights):
    g = oracle.Graph.rmat(12, 16, seed=2, weights=weights)
    off, col, w = g.csr()
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col, w)
        ops = ctx.generate_batch(seed=2, batch_id=3, n_del=700, n_ins=800, wmin=1, wmax=64)
    o_ops = g.gen_batch(seed=2, batch_id=3, n_del=700, n_ins=800, wmin=1, wmax=64)
    assert np.array_equal(ops.view(np.uint32), o_ops.view(np.uint32))


@pytest.mark.parametrize("weights", [None, "u32", "f32"])
def test_apply_batch_equals_oracle(gpu, weights):
    g = oracle.Graph.rmat(12, 16, seed=3, weights=weights)
    off, col, w = g.csr()
    rng = np.random.default_rng(1)
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col, w)
        for b in range(3):
            ops = g.gen_batch(seed=3, batch_id=b, n_del=600, n_ins=600)
            # duplicates, add-over-delete, deletes of absent pairs, re-adds of existing pairs
            extra = ops[rng.integers(0, len(ops), 100)].copy()
            extra["type"][:50] = 0
            extra["w_u32"][:50] = 9
            extra["w_f32"][:50] = 0.5
            ops = np.concatenate([ops, extra])
            ops[-1] = (1, 5, 17, 0, 0, 0)
            ctx.apply_batch(ops)
            g, touched = g.apply_batch(ops)
            o_off, o_col, o_w = g.csr()
            off2, col2, w2 = ctx.download_csr()
            assert np.array_equal(off2, o_off)
            assert np.array_equal(col2, o_col)
            if weights:
                assert np.array_equal(w2.view(np.uint32), o_w.view(np.uint32))
            # the in-adjacency follows the same batch
            c_off, c_row, c_w = csc_from_csr(o_off, o_col, o_w)
            ioff, irow, iw = ctx.download_csc()
            assert np.array_equal(ioff, c_off) and np.array_equal(irow, c_row)
            if weights:
                assert np.array_equal(iw.view(np.uint32), c_w.view(np.uint32))


@pytest.mark.parametrize("weights", [None, "u32"])
def test_apply_batch_concentrated_rows(gpu, weights):
    """A batch packed into a few rows: hundreds of deletes and adds inside one
    out-row and hundreds of adds into one in-row put more groups into a copy tile
    than the merge stages in LDS (its per-edge path); the rest of the graph takes
    the staged path in the same merge.  CSR and CSC bit-exact vs the oracle."""
    g = oracle.Graph.rmat(12, 16, seed=6, weights=weights)
    off, col, w = g.csr()
    n = g.n
    deg = np.diff(off.astype(np.int64))
    hub = int(np.argmax(deg))
    row = col[off[hub]:off[hub + 1]]
    rng = np.random.default_rng(5)
    present = np.zeros(n, bool)
    present[row] = True
    dels = rng.choice(row, size=min(len(row) // 2, 600), replace=False)
    absent = np.nonzero(~present)[0]
    adds = rng.choice(absent, size=600, replace=False)
    sink = int(np.argmin(np.bincount(col, minlength=n)))  # an in-row with few entries
    srcs = rng.choice(np.setdiff1d(np.arange(n), [sink]), size=700, replace=False)
    ops = np.zeros(len(dels) + len(adds) + len(srcs), gpu.UPDATE_DTYPE)
    k = 0
    for v in dels:
        ops[k] = (1, hub, v, 0, 0, 0); k += 1
    for v in adds:
        ops[k] = (0, hub, v, 7, 0, 0); k += 1
    for u in srcs:
        ops[k] = (0, u, sink, 3, 0, 0); k += 1
    ops = ops[rng.permutation(len(ops))]
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col, w)
        ctx.apply_batch(ops)
        g2, _ = g.apply_batch(ops)
        o_off, o_col, o_w = g2.csr()
        off2, col2, w2 = ctx.download_csr()
        assert np.array_equal(off2, o_off)
        assert np.array_equal(col2, o_col)
        if weights:
            assert np.array_equal(w2.view(np.uint32), o_w.view(np.uint32))
        c_off, c_row, c_w = csc_from_csr(o_off, o_col, o_w)
        ioff, irow, iw = ctx.download_csc()
        assert np.array_equal(ioff, c_off) and np.array_equal(irow, c_row)
        if weights:
            assert np.array_equal(iw.view(np.uint32), c_w.view(np.uint32))


def test_apply_batch_rejects_bad_ops(gpu):
    g = oracle.Graph.rmat(8, 8, seed=1)
    off, col, _ = g.csr()
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col)
        ops = np.zeros(2, gpu.UPDATE_DTYPE)
        ops[0] = (0, 1, 2, 1, 0, 0)
        ops[1] = (0, 1, g.n + 3, 1, 0, 0)
        with pytest.raises(gpu.SumincError) as e:
            ctx.apply_batch(ops)
        assert "line 1" in str(e.value)
        ops[1] = (5, 1, 2, 1, 0, 0)
        with pytest.raises(gpu.SumincError):
            ctx.apply_batch(ops)
        # the graph is unchanged after a rejected batch
        o2, c2, _ = ctx.download_csr()
        assert np.array_equal(o2, off) and np.array_equal(c2, col)


@pytest.mark.parametrize("weights", [None, "u32"])
def test_apply_batch_device_equals_host_path(gpu, weights):
    """Device-resident ops give the same graph and the same SSSP repair as host ops."""
    g = oracle.Graph.rmat(12, 16, seed=4, weights=weights)
    off, col, w = g.csr()
    with gpu.Context() as a, gpu.Context() as b:
        for ctx in (a, b):
            ctx.upload_csr(off, col, w)
            ctx.sssp_init(0)
            ctx.sssp_run()
        for i in range(3):
            ops = g.gen_batch(seed=4, batch_id=i, n_del=500, n_ins=500)
            a.apply_batch(ops)
            a.sssp_run()
            db = b.upload_batch(ops)
            b.apply_batch_device(db)
            b.sssp_run()
            db.free()
            for x, y in zip(a.download_csr(), b.download_csr()):
                assert (x is None and y is None) or np.array_equal(x, y)
            for x, y in zip(a.sssp_download(), b.sssp_download()):
                assert np.array_equal(x, y)


def test_apply_batch_device_rejects_bad_ops(gpu):
    g = oracle.Graph.rmat(8, 8, seed=1, weights="u32")
    off, col, w = g.csr()
    with gpu.Context() as ctx:
        ctx.upload_csr(off, col, w)
        ops = np.zeros(4, gpu.UPDATE_DTYPE)
        ops[0] = (0, 1, 2, 1, 0, 0)
        ops[1] = (0, 1, 3, 1, 0, 0)
        ops[2] = (0, 1, g.n + 3, 1, 0, 0)   # id out of range
        ops[3] = (0, 1, 4, 0, 0, 0)         # zero integer weight
        db = ctx.upload_batch(ops)
        with pytest.raises(gpu.SumincError) as e:
            ctx.apply_batch_device(db)
        assert "2 invalid ops" in str(e.value) and "line 2" in str(e.value)
        db.free()
        o2, c2, _ = ctx.download_csr()
        assert np.array_equal(o2, off) and np.array_equal(c2, col)
