"""ctypes handle over the oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by the cpu_baseline leg of bench.py, always as the
checker or the timed CPU baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

UPDATE_DTYPE = np.dtype(
    [("type", "<u4"), ("u", "<u4"), ("v", "<u4"), ("w_u32", "<u4"), ("w_f32", "<f4"), ("pad", "<u4")]
)

_lib = None
vp = C.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        sig = {
            "orc_philox4x32_10": (None, [vp, vp, vp]),
            "orc_num_threads": (C.c_int, []),
            "orc_set_threads": (None, [C.c_int]),
            "orc_graph_from_csr": (vp, [C.c_uint32, C.c_uint64, vp, vp, vp, vp]),
            "orc_graph_info": (None, [vp, vp, vp, vp, vp]),
            "orc_graph_copy": (None, [vp, vp, vp, vp, vp]),
            "orc_graph_free": (None, [vp]),
            "orc_rmat": (vp, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int,
                              C.c_uint32, C.c_uint32]),
            "orc_gen_batch": (None, [vp, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, C.c_int, C.c_uint32,
                                     C.c_uint32, vp]),
            "orc_apply_batch": (vp, [vp, vp, C.c_uint64, vp]),
            "orc_pr_ingress": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_int, vp, vp, vp]),
            "orc_pr_solve": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, vp, vp, C.c_int]),
            "orc_pr_amend": (None, [vp, C.c_double, C.c_int, vp, vp, vp]),
            "orc_sssp_dijkstra": (None, [vp, C.c_uint32, vp, vp]),
            "orc_sssp_ingress": (C.c_int, [vp, C.c_uint32, vp, vp, vp]),
            "orc_sssp_inc": (None, [vp, vp, C.c_uint32, vp, vp, vp]),
            "orc_leaf_clusters": (None, [vp, C.c_uint32, C.c_uint32, vp]),
            "orc_sum_structure": (None, [vp, vp, C.c_uint32, C.c_uint32, C.c_double, vp, vp]),
            "orc_sum_shortcut": (C.c_int, [vp, vp, C.c_uint32, C.c_double, C.c_double, C.c_int, vp, vp]),
            "orc_sum_pr": (None, [vp, vp, vp, C.c_double, C.c_double, C.c_int, C.c_int, vp, vp, vp, vp]),
            "orc_sum_sssp_shortcut": (None, [vp, vp, C.c_uint32, vp]),
            "orc_sum_sssp": (None, [vp, vp, vp, C.c_uint32, vp, vp, vp]),
            "orc_wpr_forward": (C.c_int, [vp, vp, vp, C.c_double, C.c_double, C.c_int, vp]),
            "orc_wpr_adjoint": (C.c_int, [vp, vp, vp, vp, C.c_double, C.c_double, C.c_int, vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(vp)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def set_threads(t: int):
    lib().orc_set_threads(int(t))


def num_threads() -> int:
    return lib().orc_num_threads()


class Graph:
    """CSR graph owned by the oracle library."""

    def __init__(self, handle):
        self.h = handle
        n, m, hu, hf = C.c_uint32(), C.c_uint64(), C.c_int(), C.c_int()
        lib().orc_graph_info(self.h, C.byref(n), C.byref(m), C.byref(hu), C.byref(hf))
        self.n, self.m = n.value, m.value
        self.wkind = 1 if hu.value else (2 if hf.value else 0)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_graph_free(self.h)
            self.h = None

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, weights=None, wmin=1, wmax=64):
        kind = {None: 0, "u32": 1, "f32": 2}[weights]
        return cls(lib().orc_rmat(scale, edge_factor, a, b, c, seed, kind, wmin, wmax))

    @classmethod
    def from_csr(cls, off, col, weights=None):
        off = np.ascontiguousarray(off, np.uint64)
        col = np.ascontiguousarray(col, np.uint32)
        wu = wf = None
        if weights is not None:
            if np.asarray(weights).dtype == np.float32:
                wf = np.ascontiguousarray(weights, np.float32)
            else:
                wu = np.ascontiguousarray(weights, np.uint32)
        return cls(lib().orc_graph_from_csr(len(off) - 1, len(col), _p(off), _p(col), _p(wu), _p(wf)))

    @classmethod
    def from_edges(cls, n, src, dst, weights=None):
        src = np.asarray(src, np.int64)
        dst = np.asarray(dst, np.int64)
        order = np.lexsort((dst, src))
        src, dst = src[order], dst[order]
        off = np.zeros(n + 1, np.uint64)
        np.add.at(off, src + 1, 1)
        off = np.cumsum(off).astype(np.uint64)
        w = None if weights is None else np.asarray(weights)[order]
        return cls.from_csr(off, dst.astype(np.uint32), w)

    def csr(self):
        off = np.zeros(self.n + 1, np.uint64)
        col = np.zeros(self.m, np.uint32)
        wu = np.zeros(self.m, np.uint32) if self.wkind == 1 else None
        wf = np.zeros(self.m, np.float32) if self.wkind == 2 else None
        lib().orc_graph_copy(self.h, _p(off), _p(col), _p(wu), _p(wf))
        return off, col, (wu if wu is not None else wf)

    def gen_batch(self, seed, batch_id, n_del, n_ins, wmin=1, wmax=64):
        ops = np.zeros(n_del + n_ins, UPDATE_DTYPE)
        lib().orc_gen_batch(self.h, seed, batch_id, n_del, n_ins, self.wkind, wmin, wmax, _p(ops))
        return ops

    def apply_batch(self, ops):
        ops = np.ascontiguousarray(ops, UPDATE_DTYPE)
        touched = np.zeros(self.n, np.uint8)
        g = Graph(lib().orc_apply_batch(self.h, _p(ops), len(ops), _p(touched)))
        return g, touched

    def pr_ingress(self, d, T, max_rounds=100, r=None, delta=None):
        init = r is None
        r = np.zeros(self.n) if r is None else np.ascontiguousarray(r, np.float64).copy()
        delta = np.zeros(self.n) if delta is None else np.ascontiguousarray(delta, np.float64).copy()
        st = np.zeros(3)
        lib().orc_pr_ingress(self.h, d, T, max_rounds, int(init), _p(r), _p(delta), _p(st))
        return r, delta, {"rounds": st[0], "edges": st[1], "residual": st[2]}

    def pr_solve(self, d, tol=1e-14, max_rounds=100000, r=None, delta=None):
        init = r is None
        r = np.zeros(self.n) if r is None else np.ascontiguousarray(r, np.float64).copy()
        delta = np.zeros(self.n) if delta is None else np.ascontiguousarray(delta, np.float64).copy()
        rounds = lib().orc_pr_solve(self.h, d, tol, max_rounds, _p(r), _p(delta), int(init))
        return r, delta, rounds

    def pr_amend(self, d, type_, r, delta, touched=None):
        r = np.ascontiguousarray(r, np.float64)
        delta = np.ascontiguousarray(delta, np.float64)
        t = None if touched is None else np.ascontiguousarray(touched, np.uint8)
        lib().orc_pr_amend(self.h, d, int(type_), _p(r), _p(delta), _p(t))

    def dijkstra(self, src=0):
        dist = np.zeros(self.n, np.uint32)
        par = np.zeros(self.n, np.uint32)
        lib().orc_sssp_dijkstra(self.h, src, _p(dist), _p(par))
        return dist, par

    def sssp_ingress(self, src=0):
        dist = np.zeros(self.n, np.uint32)
        par = np.zeros(self.n, np.uint32)
        st = np.zeros(2)
        lib().orc_sssp_ingress(self.h, src, _p(dist), _p(par), _p(st))
        return dist, par, {"rounds": st[0], "edges": st[1]}

    def sssp_inc(self, old, src, dist, parent):
        """Incremental SSSP step on this (new) graph from the state on `old`."""
        dist = np.ascontiguousarray(dist, np.uint32).copy()
        par = np.ascontiguousarray(parent, np.uint32).copy()
        st = np.zeros(3)
        lib().orc_sssp_inc(old.h, self.h, src, _p(dist), _p(par), _p(st))
        return dist, par, {"rounds": st[0], "edges": st[1], "reset": st[2]}

    # ---- supernode compression ------------------------------------------------
    def leaf_clusters(self, max_node=1024, min_node=2):
        raw = np.zeros(self.n, np.uint32)
        lib().orc_leaf_clusters(self.h, max_node, min_node, _p(raw))
        return raw

    def sum_structure(self, raw, min_node=2, max_node=1024, threshold=1.0):
        raw = np.ascontiguousarray(raw, np.uint32)
        cid = np.zeros(self.n, np.uint32)
        typ = np.zeros(self.n, np.uint8)
        lib().orc_sum_structure(self.h, _p(raw), min_node, max_node, threshold, _p(cid), _p(typ))
        return cid, typ

    def sum_shortcut(self, cid, entry, d, thr, maxit=100):
        cid = np.ascontiguousarray(cid, np.uint32)
        r = np.zeros(self.n)
        dl = np.zeros(self.n)
        steps = lib().orc_sum_shortcut(self.h, _p(cid), entry, d, thr, maxit, _p(r), _p(dl))
        return r, dl, steps

    def sum_pr(self, cid, typ, d, T, max_rounds=100000, r=None, delta=None, dirty=None):
        init = r is None
        cid = np.ascontiguousarray(cid, np.uint32)
        typ = np.ascontiguousarray(typ, np.uint8)
        r = np.zeros(self.n) if r is None else np.ascontiguousarray(r, np.float64).copy()
        delta = np.zeros(self.n) if delta is None else np.ascontiguousarray(delta, np.float64).copy()
        dv = None if dirty is None else np.ascontiguousarray(dirty, np.uint8)
        st = np.zeros(4)
        lib().orc_sum_pr(self.h, _p(cid), _p(typ), d, T, max_rounds, int(init), _p(dv), _p(r), _p(delta), _p(st))
        return r, delta, {"sketch_rounds": st[0], "plain_rounds": st[1], "edges": st[2], "residual": st[3]}

    def sum_sssp_shortcut(self, cid, entry):
        cid = np.ascontiguousarray(cid, np.uint32)
        dist = np.zeros(self.n, np.uint32)
        lib().orc_sum_sssp_shortcut(self.h, _p(cid), entry, _p(dist))
        return dist

    def sum_sssp(self, cid, typ, src=0):
        cid = np.ascontiguousarray(cid, np.uint32)
        typ = np.ascontiguousarray(typ, np.uint8)
        dist = np.zeros(self.n, np.uint32)
        par = np.zeros(self.n, np.uint32)
        st = np.zeros(2)
        lib().orc_sum_sssp(self.h, _p(cid), _p(typ), src, _p(dist), _p(par), _p(st))
        return dist, par, {"rounds": st[0], "edges": st[1]}

    # ---- weighted personalized PageRank + adjoint (FP64) -----------------------
    def wpr_forward(self, w, p, d, tol=1e-15, max_rounds=100000):
        w = np.ascontiguousarray(w, np.float32)
        p = np.ascontiguousarray(p, np.float64)
        r = np.zeros(self.n)
        rounds = lib().orc_wpr_forward(self.h, _p(w), _p(p), d, tol, max_rounds, _p(r))
        return r, rounds

    def wpr_adjoint(self, w, r, c, d, tol=1e-15, max_rounds=100000):
        w = np.ascontiguousarray(w, np.float32)
        r = np.ascontiguousarray(r, np.float64)
        c = np.ascontiguousarray(c, np.float64)
        lam = np.zeros(self.n)
        gp = np.zeros(self.n)
        gw = np.zeros(len(w))
        lib().orc_wpr_adjoint(self.h, _p(w), _p(r), _p(c), d, tol, max_rounds, _p(lam), _p(gp), _p(gw))
        return lam, gp, gw
