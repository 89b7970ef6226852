"""Python handle over one libsuminc context (one device, one stream).

This is plumbing for tests, the benchmark and the autograd wrapper
(suminc_amd/wpr.py); every call goes through the C ABI of include/suminc.h.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L


def _buf(x, dtype, count):
    """Pointer to `count` elements of `dtype` in a numpy array or a device tensor."""
    if hasattr(x, "data_ptr"):
        if x.numel() != count or x.element_size() != np.dtype(dtype).itemsize or not x.is_contiguous():
            raise ValueError(f"expected a contiguous tensor of {count} {np.dtype(dtype).name} values")
        return C.c_void_p(x.data_ptr())
    if not isinstance(x, np.ndarray) or x.dtype != dtype or x.size != count or not x.flags["C_CONTIGUOUS"]:
        raise ValueError(f"expected a contiguous numpy array of {count} {np.dtype(dtype).name} values")
    return L.ptr(x)


class Context:
    def __init__(self, device: int = 0, **params):
        self._lib = L.lib()
        p = L.Params()
        self._lib.suminc_default_params(C.byref(p))
        for k, v in params.items():
            if not hasattr(p, k):
                raise TypeError(f"unknown parameter {k!r}")
            setattr(p, k, v)
        h = C.c_void_p()
        rc = self._lib.suminc_ctx_create(device, C.byref(p), C.byref(h))
        if rc != L.OK:
            raise L.SumincError(rc, "suminc_ctx_create failed (is a GPU visible?)")
        self._h = h
        self.params = p

    # ---- plumbing -------------------------------------------------------------
    def _check(self, rc: int):
        if rc != L.OK:
            msg = self._lib.suminc_last_error(self._h)
            raise L.SumincError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.suminc_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_params(self, **params):
        for k, v in params.items():
            if not hasattr(self.params, k):
                raise TypeError(f"unknown parameter {k!r}")
            setattr(self.params, k, v)
        self._check(self._lib.suminc_set_params(self._h, C.byref(self.params)))

    def sync(self):
        self._check(self._lib.suminc_sync(self._h))

    @property
    def stream(self) -> int:
        s = C.c_void_p()
        self._check(self._lib.suminc_get_stream(self._h, C.byref(s)))
        return s.value or 0

    def stats(self) -> dict:
        s = L.Stats()
        self._check(self._lib.suminc_get_stats(self._h, C.byref(s)))
        return s.as_dict()

    def profile_enable(self, on: bool = True):
        self._check(self._lib.suminc_profile_enable(self._h, int(on)))

    def profile(self, kernel: int) -> dict:
        p = L.KernelProfile()
        self._check(self._lib.suminc_profile_get(self._h, kernel, C.byref(p)))
        return p.as_dict()

    # ---- graph ----------------------------------------------------------------
    def generate_rmat(self, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, weights=None, wmin=1, wmax=64):
        kind = {None: L.W_NONE, "u32": L.W_U32, "f32": L.W_F32}[weights]
        self._check(self._lib.suminc_generate_rmat(self._h, scale, edge_factor, a, b, c, seed, kind, wmin, wmax))

    def load_edge_file(self, path, weights=None, directed=True) -> dict:
        """Reads an edge file ("u v [w]" lines, the reference's efile) on the host and
        uploads it; returns n, m and the number of lines that did not parse."""
        from .textio import EdgeFile

        with EdgeFile(path, weights=weights, directed=directed) as ef:
            ef.upload(self)
            return {"n": ef.n, "m": ef.m, "skipped_lines": ef.skipped_lines}

    def upload_csr(self, offsets, cols, weights=None):
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        col = np.ascontiguousarray(cols, dtype=np.uint32)
        kind, w = L.W_NONE, None
        if weights is not None:
            w = np.ascontiguousarray(weights)
            if w.dtype == np.float32:
                kind = L.W_F32
            else:
                w = w.astype(np.uint32, copy=False)
                kind = L.W_U32
        n = len(off) - 1
        self._check(self._lib.suminc_upload_csr(self._h, n, len(col), L.ptr(off), L.ptr(col), kind, L.ptr(w)))

    def graph_info(self):
        n, m, k = C.c_uint32(), C.c_uint64(), C.c_int()
        self._check(self._lib.suminc_graph_info(self._h, C.byref(n), C.byref(m), C.byref(k)))
        return n.value, m.value, k.value

    def download_csr(self):
        n, m, kind = self.graph_info()
        off = np.zeros(n + 1, np.uint64)
        col = np.zeros(m, np.uint32)
        w = None
        if kind == L.W_U32:
            w = np.zeros(m, np.uint32)
        elif kind == L.W_F32:
            w = np.zeros(m, np.float32)
        self._check(self._lib.suminc_download_csr(self._h, L.ptr(off), L.ptr(col), L.ptr(w)))
        return off, col, w

    def download_csc(self):
        n, m, kind = self.graph_info()
        off = np.zeros(n + 1, np.uint64)
        row = np.zeros(m, np.uint32)
        w = None
        if kind == L.W_U32:
            w = np.zeros(m, np.uint32)
        elif kind == L.W_F32:
            w = np.zeros(m, np.float32)
        self._check(self._lib.suminc_download_csc(self._h, L.ptr(off), L.ptr(row), L.ptr(w)))
        return off, row, w

    # ---- batches --------------------------------------------------------------
    def generate_batch(self, seed, batch_id, n_del, n_ins, wmin=1, wmax=64):
        ops = np.zeros(n_del + n_ins, L.UPDATE_DTYPE)
        self._check(self._lib.suminc_generate_batch(self._h, seed, batch_id, n_del, n_ins, wmin, wmax, L.ptr(ops)))
        return ops

    def apply_batch(self, ops):
        ops = np.ascontiguousarray(ops, dtype=L.UPDATE_DTYPE)
        self._check(self._lib.suminc_apply_batch(self._h, L.ptr(ops), len(ops)))
        return self.stats()

    def upload_batch(self, ops) -> "DeviceBatch":
        """Copies a batch into device memory (for apply_batch_device)."""
        ops = np.ascontiguousarray(ops, dtype=L.UPDATE_DTYPE)
        p = C.c_void_p()
        self._check(self._lib.suminc_device_alloc(self._h, ops.nbytes, C.byref(p)))
        self._check(self._lib.suminc_copy_to_device(self._h, p, L.ptr(ops), ops.nbytes))
        return DeviceBatch(self, p, len(ops))

    def device_array(self, arr) -> "DeviceArray":
        """Copies a numpy array into device memory of this context's GPU."""
        return DeviceArray(self, arr)

    def apply_batch_device(self, batch: "DeviceBatch"):
        self._check(self._lib.suminc_apply_batch_device(self._h, batch.ptr, batch.n_ops))
        return self.stats()

    # ---- PageRank -------------------------------------------------------------
    def pr_init(self):
        self._check(self._lib.suminc_pr_init(self._h))

    def pr_run(self, allow_not_converged=False) -> dict:
        s = L.Stats()
        rc = self._lib.suminc_pr_run(self._h, C.byref(s))
        if not (allow_not_converged and rc == L.E_NOT_CONVERGED):
            self._check(rc)
        return s.as_dict()

    def pr_step(self) -> float:
        r = C.c_double()
        self._check(self._lib.suminc_pr_step(self._h, C.byref(r)))
        return r.value

    def pr_download(self):
        n = self.graph_info()[0]
        rank = np.zeros(n, np.float64)
        delta = np.zeros(n, np.float64)
        self._check(self._lib.suminc_pr_download(self._h, L.ptr(rank), L.ptr(delta)))
        return rank, delta

    def pr_upload(self, rank, delta):
        r = np.ascontiguousarray(rank, np.float64)
        d = np.ascontiguousarray(delta, np.float64)
        self._check(self._lib.suminc_pr_upload(self._h, L.ptr(r), L.ptr(d)))

    # ---- supernode sketch -----------------------------------------------------
    def cluster_leaves(self, max_node_num=1024, min_node_num=2):
        n = self.graph_info()[0]
        out = np.zeros(n, np.uint32)
        self._check(self._lib.suminc_cluster_leaves(self._h, max_node_num, min_node_num, L.ptr(out)))
        return out

    def set_clusters(self, cluster_of):
        c = np.ascontiguousarray(cluster_of, np.uint32)
        self._check(self._lib.suminc_set_clusters(self._h, L.ptr(c)))

    def sum_build(self, min_node_num=2, max_node_num=1024, compress_threshold=1.0):
        self._check(self._lib.suminc_sum_build(self._h, min_node_num, max_node_num, compress_threshold))
        return self.sum_info()

    def sum_info(self) -> dict:
        i = L.SketchInfo()
        self._check(self._lib.suminc_sum_info(self._h, C.byref(i)))
        return i.as_dict()

    def sum_download(self):
        n = self.graph_info()[0]
        cid = np.zeros(n, np.uint32)
        typ = np.zeros(n, np.uint8)
        self._check(self._lib.suminc_sum_download(self._h, L.ptr(cid), L.ptr(typ)))
        return cid, typ

    def sum_pr_shortcut(self, entry):
        cnt, nex = C.c_uint32(), C.c_uint32()
        self._check(self._lib.suminc_sum_pr_shortcut(self._h, entry, 0, None, None, None, C.byref(cnt), C.byref(nex)))
        mem = np.zeros(cnt.value, np.uint32)
        reach = np.zeros(cnt.value)
        resid = np.zeros(cnt.value)
        self._check(self._lib.suminc_sum_pr_shortcut(self._h, entry, cnt.value, L.ptr(mem), L.ptr(reach), L.ptr(resid),
                                                     C.byref(cnt), C.byref(nex)))
        return mem, reach, resid, nex.value

    def sum_pr_run(self) -> dict:
        s = L.Stats()
        self._check(self._lib.suminc_sum_pr_run(self._h, C.byref(s)))
        return s.as_dict()

    def sum_sssp_shortcut(self, entry):
        cnt, nex = C.c_uint32(), C.c_uint32()
        self._check(self._lib.suminc_sum_sssp_shortcut(self._h, entry, 0, None, None, None, C.byref(cnt),
                                                       C.byref(nex)))
        mem = np.zeros(cnt.value, np.uint32)
        dist = np.zeros(cnt.value, np.uint32)
        par = np.zeros(cnt.value, np.uint32)
        self._check(self._lib.suminc_sum_sssp_shortcut(self._h, entry, cnt.value, L.ptr(mem), L.ptr(dist), L.ptr(par),
                                                       C.byref(cnt), C.byref(nex)))
        return mem, dist, par, nex.value

    def sum_sssp_run(self) -> dict:
        s = L.Stats()
        self._check(self._lib.suminc_sum_sssp_run(self._h, C.byref(s)))
        return s.as_dict()

    # ---- the reference compressor's sketch, phase by phase --------------------
    def compress_mirror(self, cluster_of, min_node_num=2, max_node_num=1024, compress_threshold=1.0,
                        mirror_k=4.0) -> dict:
        """The compressor with mirrors; its products become the context's sketch."""
        c = np.ascontiguousarray(cluster_of, np.uint32)
        info = L.CompressInfo()
        self._check(self._lib.suminc_compress_mirror(self._h, L.ptr(c), min_node_num, max_node_num,
                                                     float(compress_threshold), float(mirror_k), C.byref(info)))
        self._compress_info = info.as_dict()
        return self._compress_info

    def compress_mirror_save(self, path):
        """The accepted clusters of the last compress_mirror, in the reference's cluster cache layout."""
        self._check(self._lib.suminc_compress_mirror_save(self._h, os.fsencode(os.fspath(path))))

    def compress_mirror_load(self, path) -> dict:
        """compress_mirror's products from a cluster cache (no acceptance pass)."""
        info = L.CompressInfo()
        self._check(self._lib.suminc_compress_mirror_load(self._h, os.fsencode(os.fspath(path)), C.byref(info)))
        self._compress_info = info.as_dict()
        return self._compress_info

    def sketch_download(self) -> dict:
        """The last compress_mirror's products as numpy arrays (suminc_sketch_csr layout)."""
        i = self._compress_info
        n = self.graph_info()[0]
        a = i["all_node_num"]
        out = {"n": n, "all_node_num": a,
               "node_type": np.zeros(n, np.uint8), "mirror_master": np.zeros(a - n, np.uint32),
               "cluster_of": np.zeros(a, np.uint32),
               "is_off": np.zeros(a + 1, np.uint64), "is_col": np.zeros(i["is_edges"], np.uint32),
               "is_w": np.zeros(i["is_edges"], np.float32),
               "ib_off": np.zeros(n + 1, np.uint64), "ib_col": np.zeros(i["ib_edges"], np.uint32),
               "sync_off": np.zeros(a + 1, np.uint64), "sync_col": np.zeros(i["sync_edges"], np.uint32),
               "iv_off": np.zeros(a + 1, np.uint64), "iv_col": np.zeros(i["inner_value_entries"], np.uint32),
               "iv_w": np.zeros(i["inner_value_entries"], np.float32),
               "id_off": np.zeros(a + 1, np.uint64), "id_col": np.zeros(i["inner_delta_entries"], np.uint32),
               "id_w": np.zeros(i["inner_delta_entries"], np.float32),
               "sub_off": np.zeros(a + 1, np.uint64), "sub_col": np.zeros(i["subgraph_edges"], np.uint32)}
        b = L.SketchBuffers()
        for name, _ in b._fields_:
            setattr(b, name, out[name].ctypes.data if out[name].size else None)
        self._check(self._lib.suminc_sketch_download(self._h, C.byref(b)))
        return out

    def upload_sketch(self, sk: dict):
        """Uploads a sketch in the reference compressor's layout (dict of arrays as
        sketch_download returns)."""
        arrs = {}
        dt = {"node_type": np.uint8, "mirror_master": np.uint32, "is_off": np.uint64, "is_col": np.uint32,
              "is_w": np.float32, "ib_off": np.uint64, "ib_col": np.uint32, "sync_off": np.uint64,
              "sync_col": np.uint32, "iv_off": np.uint64, "iv_col": np.uint32, "iv_w": np.float32,
              "id_off": np.uint64, "id_col": np.uint32, "id_w": np.float32, "sub_off": np.uint64,
              "sub_col": np.uint32}
        s = L.SketchCsr()
        s.n = int(sk["n"])
        s.all_node_num = int(sk["all_node_num"])
        for name, t in dt.items():
            if sk.get(name) is None:  # the subgraph rows are optional
                continue
            arrs[name] = np.ascontiguousarray(sk[name], t)
            setattr(s, name, arrs[name].ctypes.data if arrs[name].size else None)
        self._check(self._lib.suminc_upload_sketch(self._h, C.byref(s)))

    def pr_step_phase(self, phase) -> float:
        r = C.c_double()
        self._check(self._lib.suminc_pr_step_phase(self._h, int(phase), C.byref(r)))
        return r.value

    def pr_correct(self):
        self._check(self._lib.suminc_pr_correct(self._h))

    def pr_precompute(self):
        self._check(self._lib.suminc_pr_precompute(self._h))

    def sketch_pr_run(self) -> dict:
        s = L.Stats()
        self._check(self._lib.suminc_sketch_pr_run(self._h, C.byref(s)))
        return s.as_dict()

    # ---- weighted personalized PageRank + adjoint ------------------------------
    # Arrays may be numpy arrays (host) or torch tensors on this context's GPU.
    # The backward reads the rank and weights of the last forward on the context;
    # wpr_token counts forwards and weight changes, so a caller holding an older
    # forward (the autograd wrapper) can tell that it must run its forward again.
    wpr_token = 0

    def wpr_set_weights(self, w):
        self.wpr_token += 1
        self._check(self._lib.suminc_wpr_set_weights(self._h, _buf(w, np.float32, self.graph_info()[1])))

    def wpr_forward(self, p, out=None) -> tuple:
        n = self.graph_info()[0]
        r = np.zeros(n, np.float32) if out is None else out
        s = L.Stats()
        self.wpr_token += 1
        self._check(self._lib.suminc_wpr_forward(self._h, _buf(p, np.float32, n), _buf(r, np.float32, n),
                                                 C.byref(s)))
        return r, s.as_dict()

    def wpr_backward(self, c, grad_p=None, grad_w=None) -> tuple:
        n, m, _ = self.graph_info()
        gp = np.zeros(n, np.float32) if grad_p is None else grad_p
        gw = np.zeros(m, np.float32) if grad_w is None else grad_w
        s = L.Stats()
        self._check(self._lib.suminc_wpr_backward(self._h, _buf(c, np.float32, n), _buf(gp, np.float32, n),
                                                  _buf(gw, np.float32, m), C.byref(s)))
        return gp, gw, s.as_dict()

    def wpr_download(self, adjoint=False):
        """FP64 r of the last forward and, with adjoint=True, lambda of the last backward."""
        n = self.graph_info()[0]
        r = np.zeros(n)
        lam = np.zeros(n) if adjoint else None
        self._check(self._lib.suminc_wpr_download(self._h, L.ptr(r), L.ptr(lam)))
        return r, lam

    # ---- SSSP -----------------------------------------------------------------
    def sssp_init(self, source=0):
        self._check(self._lib.suminc_sssp_init(self._h, source))

    def traversal_init(self, app, source=0):
        """SSSP, BFS or SSWP state (APP_SSSP / APP_BFS / APP_SSWP) for sssp_run / sssp_step /
        sssp_download and the batch repair."""
        self._check(self._lib.suminc_traversal_init(self._h, int(app), source))

    def php_run(self, source=0, damping=float(np.float32(0.8))):
        """PHP values (FP64) from `source`; the reference's d is 0.80f."""
        n = self.graph_info()[0]
        out = np.zeros(n)
        s = L.Stats()
        self._check(self._lib.suminc_php_run(self._h, source, float(damping), L.ptr(out), C.byref(s)))
        return out, s.as_dict()

    def sssp_run(self) -> dict:
        s = L.Stats()
        self._check(self._lib.suminc_sssp_run(self._h, C.byref(s)))
        return s.as_dict()

    def sssp_step(self) -> int:
        """One min-plus round; returns the size of the frontier it produced."""
        f = C.c_uint32()
        self._check(self._lib.suminc_sssp_step(self._h, C.byref(f)))
        return f.value

    def sssp_download(self):
        n = self.graph_info()[0]
        dist = np.zeros(n, np.uint32)
        parent = np.zeros(n, np.uint32)
        self._check(self._lib.suminc_sssp_download(self._h, L.ptr(dist), L.ptr(parent)))
        return dist, parent


class DeviceArray:
    """A numpy array copied into device memory of a context's GPU (a caller-owned
    suminc_device_alloc buffer).  It exposes the tensor protocol `_buf` reads, so it
    can be passed wherever the ABI accepts device memory; `set` rewrites single
    elements from the host."""

    def __init__(self, ctx: Context, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        self._ctx = ctx
        self.dtype = arr.dtype
        self.size = arr.size
        self.ptr = C.c_void_p()
        ctx._check(ctx._lib.suminc_device_alloc(ctx._h, max(arr.nbytes, 1), C.byref(self.ptr)))
        ctx._check(ctx._lib.suminc_copy_to_device(ctx._h, self.ptr, L.ptr(arr), arr.nbytes))

    def data_ptr(self) -> int:
        return self.ptr.value

    def numel(self) -> int:
        return self.size

    def element_size(self) -> int:
        return self.dtype.itemsize

    def is_contiguous(self) -> bool:
        return True

    def download(self) -> np.ndarray:
        out = np.zeros(self.size, self.dtype)
        self._ctx._check(self._ctx._lib.suminc_copy_from_device(self._ctx._h, L.ptr(out), self.ptr, out.nbytes))
        return out

    def set(self, index: int, value):
        v = np.array([value], dtype=self.dtype)
        dst = C.c_void_p(self.ptr.value + int(index) * self.dtype.itemsize)
        self._ctx._check(self._ctx._lib.suminc_copy_to_device(self._ctx._h, dst, L.ptr(v), v.nbytes))

    def free(self):
        if self.ptr is not None and self._ctx._h:
            self._ctx._check(self._ctx._lib.suminc_device_free(self._ctx._h, self.ptr))
        self.ptr = None


class DeviceBatch:
    """An update batch resident in device memory, freed with its context's device."""

    def __init__(self, ctx: Context, ptr, n_ops: int):
        self._ctx = ctx
        self.ptr = ptr
        self.n_ops = n_ops

    def free(self):
        if self.ptr is not None and self._ctx._h:
            self._ctx._check(self._ctx._lib.suminc_device_free(self._ctx._h, self.ptr))
        self.ptr = None
