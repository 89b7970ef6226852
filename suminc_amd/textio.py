"""The reference's text formats through libsuminc's host parsers (include/suminc.h).

* edge files     "u v [w]" per line (grape/fragment/e_fragment_loader.h:107-170)
* update files   "a u v [w]" / "d u v" per line (grape/fragment/inc_fragment_builder.h:29-72)
* cluster files  "<size> <id> ... <id>" per cluster, named "<efile>.c_<max_node_num>"
                 (grape/fragment/compressor_base.h:3153-3250)
* binary caches  of a parsed edge list and of update records, under the reference's
                 serialization path rule (grape/fragment/loader.h:30-61)

The parsing runs in libsuminc (multi-threaded C++); these wrappers only move the
results into numpy arrays.  None of them needs a GPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L

_WEIGHT_KINDS = {None: L.W_NONE, "none": L.W_NONE, "u32": L.W_U32, "f32": L.W_F32}


def _kind(weights) -> int:
    if isinstance(weights, int):
        return weights
    return _WEIGHT_KINDS[weights]


def _check(rc: int) -> None:
    if rc != L.OK:
        msg = L.lib().suminc_last_error(None)
        raise L.SumincError(rc, msg.decode() if msg else "")


def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


class EdgeFile:
    """An edge file parsed into a CSR with sorted rows (first line of a repeated pair
    kept).  n = largest vertex id + 1.  cache=True reads the binary cache written by
    save() instead of a text file."""

    def __init__(self, path, weights=None, directed=True, cache=False):
        lib = L.lib()
        h = C.c_void_p()
        if cache:
            _check(lib.suminc_edges_load(_path(path), C.byref(h)))
        else:
            _check(lib.suminc_edges_read_tsv(_path(path), _kind(weights), 1 if directed else 0, C.byref(h)))
        self._h = h
        n, m, wk, skipped = C.c_uint32(), C.c_uint64(), C.c_int(), C.c_uint64()
        _check(lib.suminc_edges_info(h, C.byref(n), C.byref(m), C.byref(wk), C.byref(skipped)))
        self.n, self.m, self.weight_kind, self.skipped_lines = n.value, m.value, wk.value, skipped.value

    def csr(self):
        """(offsets u64[n+1], cols u32[m], weights u32/f32[m] or None)."""
        off = np.zeros(self.n + 1, np.uint64)
        col = np.zeros(self.m, np.uint32)
        w = None
        if self.weight_kind == L.W_U32:
            w = np.zeros(self.m, np.uint32)
        elif self.weight_kind == L.W_F32:
            w = np.zeros(self.m, np.float32)
        _check(L.lib().suminc_edges_copy(self._h, L.ptr(off), L.ptr(col), L.ptr(w)))
        return off, col, w

    def save(self, path) -> None:
        """The binary cache of this edge list (suminc_edges_save)."""
        _check(L.lib().suminc_edges_save(self._h, _path(path)))

    def upload(self, ctx) -> None:
        """suminc_edges_upload: the parsed CSR becomes the context's graph."""
        ctx._check(L.lib().suminc_edges_upload(ctx._h, self._h))

    def close(self) -> None:
        if self._h:
            L.lib().suminc_edges_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_updates(path, weights=None, directed=True) -> np.ndarray:
    """Update records (UPDATE_DTYPE) ready for Context.apply_batch."""
    lib = L.lib()
    cnt = C.c_uint64()
    kind = _kind(weights)
    _check(lib.suminc_updates_read_tsv(_path(path), kind, 1 if directed else 0, None, 0, C.byref(cnt)))
    ops = np.zeros(cnt.value, L.UPDATE_DTYPE)
    _check(lib.suminc_updates_read_tsv(_path(path), kind, 1 if directed else 0, L.ptr(ops), cnt.value, C.byref(cnt)))
    return ops


def save_updates(path, ops) -> None:
    """The binary cache of update records (suminc_updates_save)."""
    a = np.ascontiguousarray(ops, L.UPDATE_DTYPE)
    _check(L.lib().suminc_updates_save(_path(path), L.ptr(a), len(a)))


def load_updates(path) -> np.ndarray:
    """Update records from a binary cache (suminc_updates_load)."""
    lib = L.lib()
    cnt = C.c_uint64()
    _check(lib.suminc_updates_load(_path(path), None, 0, C.byref(cnt)))
    ops = np.zeros(cnt.value, L.UPDATE_DTYPE)
    _check(lib.suminc_updates_load(_path(path), L.ptr(ops), cnt.value, C.byref(cnt)))
    return ops


def serialization_path(prefix, efile, vfile="", workers=1, fid=0) -> str:
    """The reference's cache path "<prefix>/<efile + vfile + workers, '/' -> '_'>/frag_<fid>.s"."""
    n = C.c_uint64()
    lib = L.lib()
    _check(lib.suminc_serialization_path(_path(prefix), _path(efile), _path(vfile), workers, fid, None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    _check(lib.suminc_serialization_path(_path(prefix), _path(efile), _path(vfile), workers, fid, buf, n.value + 1,
                                         C.byref(n)))
    return os.fsdecode(buf.value)


def read_clusters(path, n: int, min_node_num: int = 2):
    """(cluster_of u32[n] with INF_U32 for unclustered vertices, number of clusters)."""
    out = np.zeros(n, np.uint32)
    k = C.c_uint32()
    _check(L.lib().suminc_clusters_read(_path(path), n, min_node_num, L.ptr(out), C.byref(k)))
    return out, k.value


def write_clusters(path, cluster_of: np.ndarray) -> None:
    a = np.ascontiguousarray(cluster_of, dtype=np.uint32)
    _check(L.lib().suminc_clusters_write(_path(path), a.size, L.ptr(a)))


def cluster_lpa(offsets, cols, max_node_num=1024, min_node_num=2, rounds=10, max_degree=4096):
    """Community clustering of a CSR (suminc_cluster_lpa): (cluster_of u32[n], count)."""
    off = np.ascontiguousarray(offsets, np.uint64)
    col = np.ascontiguousarray(cols, np.uint32)
    n = len(off) - 1
    out = np.zeros(n, np.uint32)
    k = C.c_uint32()
    _check(L.lib().suminc_cluster_lpa(n, L.ptr(off), L.ptr(col), max_node_num, min_node_num, rounds, max_degree,
                                      L.ptr(out), C.byref(k)))
    return out, k.value


def cluster_file_path(efile: str, max_node_num: int) -> str:
    """The cluster file the reference's compressor reads for an edge file:
    efile + ".c_" + max_node_num, with "_w.", "_ud.", "_w1." and ".random." folded
    to "." until none is left (compressor_base.h:3173-3190)."""
    path = f"{efile}.c_{max_node_num}"
    keys = ["_w.", "_ud.", "_w1.", ".random."]
    while True:
        changed = False
        for key in keys:
            i = path.rfind(key)
            if i != -1:
                path = path[:i] + "." + path[i + len(key):]
                changed = True
        if not changed:
            return path
