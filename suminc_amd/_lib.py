"""ctypes binding of libsuminc.so (include/suminc.h).

The shared library is built in-tree by suminc_amd/build.py.  Loading fails
loudly when it is missing: there is no CPU fallback for the device path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libsuminc.so"
HEADER = Path(__file__).resolve().parent.parent / "include" / "suminc.h"

OK = 0
E_INVALID = -1
E_HIP = -2
E_NOMEM = -3
E_STATE = -4
E_TOO_LARGE = -5
E_NOT_CONVERGED = -6
E_NUMERIC = -7
ERROR_NAMES = {
    E_INVALID: "SUMINC_E_INVALID",
    E_HIP: "SUMINC_E_HIP",
    E_NOMEM: "SUMINC_E_NOMEM",
    E_STATE: "SUMINC_E_STATE",
    E_TOO_LARGE: "SUMINC_E_TOO_LARGE",
    E_NOT_CONVERGED: "SUMINC_E_NOT_CONVERGED",
    E_NUMERIC: "SUMINC_E_NUMERIC",
}

INF_U32 = 0xFFFFFFFF
PR_FP64 = 0
PR_FP32_KAHAN = 1
W_NONE, W_U32, W_F32 = 0, 1, 2
OP_ADD, OP_DEL = 0, 1

UPDATE_DTYPE = np.dtype(
    [("type", "<u4"), ("u", "<u4"), ("v", "<u4"), ("w_u32", "<u4"), ("w_f32", "<f4"), ("pad", "<u4")]
)


class Params(C.Structure):
    _fields_ = [
        ("damping", C.c_double),
        ("termcheck_threshold", C.c_double),
        ("max_rounds", C.c_int32),
        ("pr_precision", C.c_int32),
        ("sssp_source", C.c_uint32),
        ("pull_threshold_div", C.c_int32),
        ("shortcut_max_rounds", C.c_int32),
        ("pull_band_kib", C.c_int32),
        ("pull_compact_kib", C.c_int32),
        ("pull_tile_kib", C.c_int32),
        ("pull_bin_kib", C.c_int32),
        ("pull_bin_direct", C.c_int32),
        ("reserved", C.c_int32 * 2),
    ]


class SketchInfo(C.Structure):
    _fields_ = [
        ("clusters", C.c_uint64),
        ("raw_clusters", C.c_uint64),
        ("entries", C.c_uint64),
        ("clustered_vertices", C.c_uint64),
        ("node_type_count", C.c_uint64 * 5),
        ("sketch_edges", C.c_uint64),
        ("internal_edges", C.c_uint64),
        ("shortcut_slots", C.c_uint64),
        ("max_cluster", C.c_uint64),
        ("dissolved_clusters", C.c_uint64),
        ("reserved", C.c_uint64 * 7),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("reserved", "node_type_count")}
        d["node_type_count"] = list(self.node_type_count)
        return d


class Stats(C.Structure):
    _fields_ = [
        ("rounds", C.c_double),
        ("edges_traversed", C.c_double),
        ("residual", C.c_double),
        ("kernel_ms", C.c_double),
        ("push_rounds", C.c_double),
        ("pull_rounds", C.c_double),
        ("mass_defect", C.c_double),
        ("mass_total", C.c_double),
        ("reset_vertices", C.c_double),
        ("batch_groups", C.c_double),
        ("touched_rows", C.c_double),
        ("reseed_edges", C.c_double),
        ("shortcut_edges", C.c_double),
        ("sketch_rounds", C.c_double),
        ("sketch_ms", C.c_double),
        ("residual_rounds", C.c_double),
        ("precompute_edges", C.c_double),
        ("reserved", C.c_double * 7),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class SketchCsr(C.Structure):
    """suminc_sketch_csr: the reference compressor's products (pointers into numpy
    arrays the caller keeps alive for the call)."""
    _fields_ = [
        ("n", C.c_uint32),
        ("all_node_num", C.c_uint32),
        ("node_type", C.c_void_p),
        ("mirror_master", C.c_void_p),
        ("is_off", C.c_void_p), ("is_col", C.c_void_p), ("is_w", C.c_void_p),
        ("ib_off", C.c_void_p), ("ib_col", C.c_void_p),
        ("sync_off", C.c_void_p), ("sync_col", C.c_void_p),
        ("iv_off", C.c_void_p), ("iv_col", C.c_void_p), ("iv_w", C.c_void_p),
        ("id_off", C.c_void_p), ("id_col", C.c_void_p), ("id_w", C.c_void_p),
        ("sub_off", C.c_void_p), ("sub_col", C.c_void_p),
    ]


PHASE_PLAIN, PHASE_SKETCH = 0, 1
APP_SSSP, APP_BFS, APP_SSWP = 0, 1, 2


class CompressInfo(C.Structure):
    _fields_ = [
        ("clusters", C.c_uint64),
        ("all_node_num", C.c_uint64),
        ("supernodes", C.c_uint64),
        ("in_mirrors", C.c_uint64),
        ("out_mirrors", C.c_uint64),
        ("absorbed", C.c_uint64),
        ("node_type_count", C.c_uint64 * 7),
        ("is_edges", C.c_uint64),
        ("ib_edges", C.c_uint64),
        ("sync_edges", C.c_uint64),
        ("subgraph_edges", C.c_uint64),
        ("inner_value_entries", C.c_uint64),
        ("inner_delta_entries", C.c_uint64),
        ("shortcut_work", C.c_uint64),
        ("reserved", C.c_uint64 * 8),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("reserved", "node_type_count")}
        d["node_type_count"] = list(self.node_type_count)
        return d


class SketchBuffers(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in (
        "node_type", "mirror_master", "cluster_of", "is_off", "is_col", "is_w", "ib_off", "ib_col", "sync_off",
        "sync_col", "iv_off", "iv_col", "iv_w", "id_off", "id_col", "id_w", "sub_off", "sub_col")]


class KernelProfile(C.Structure):
    _fields_ = [
        ("launches", C.c_double),
        ("ms", C.c_double),
        ("alg_bytes", C.c_double),
        ("edges", C.c_double),
        ("vertices", C.c_double),
        ("reserved", C.c_double * 3),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class SumincError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")
        self.code = code


_lib = None

vp = C.c_void_p
u32, u64, i32, f64 = C.c_uint32, C.c_uint64, C.c_int, C.c_double

SIGNATURES = {
    "suminc_default_params": (None, [C.POINTER(Params)]),
    "suminc_ctx_create": (i32, [i32, C.POINTER(Params), C.POINTER(vp)]),
    "suminc_ctx_destroy": (i32, [vp]),
    "suminc_set_params": (i32, [vp, C.POINTER(Params)]),
    "suminc_last_error": (C.c_char_p, [vp]),
    "suminc_sync": (i32, [vp]),
    "suminc_get_stream": (i32, [vp, C.POINTER(vp)]),
    "suminc_upload_csr": (i32, [vp, u32, u64, vp, vp, i32, vp]),
    "suminc_generate_rmat": (i32, [vp, i32, i32, f64, f64, f64, u64, i32, u32, u32]),
    "suminc_graph_info": (i32, [vp, C.POINTER(u32), C.POINTER(u64), C.POINTER(i32)]),
    "suminc_download_csr": (i32, [vp, vp, vp, vp]),
    "suminc_download_csc": (i32, [vp, vp, vp, vp]),
    "suminc_generate_batch": (i32, [vp, u64, u32, u64, u64, u32, u32, vp]),
    "suminc_apply_batch": (i32, [vp, vp, u64]),
    "suminc_apply_batch_device": (i32, [vp, vp, u64]),
    "suminc_device_alloc": (i32, [vp, u64, C.POINTER(vp)]),
    "suminc_device_free": (i32, [vp, vp]),
    "suminc_copy_to_device": (i32, [vp, vp, vp, u64]),
    "suminc_copy_from_device": (i32, [vp, vp, vp, u64]),
    "suminc_pr_init": (i32, [vp]),
    "suminc_pr_run": (i32, [vp, C.POINTER(Stats)]),
    "suminc_pr_step": (i32, [vp, C.POINTER(f64)]),
    "suminc_pr_download": (i32, [vp, vp, vp]),
    "suminc_pr_upload": (i32, [vp, vp, vp]),
    "suminc_sssp_init": (i32, [vp, u32]),
    "suminc_sssp_run": (i32, [vp, C.POINTER(Stats)]),
    "suminc_sssp_download": (i32, [vp, vp, vp]),
    "suminc_sssp_step": (i32, [vp, C.POINTER(u32)]),
    "suminc_upload_sketch": (i32, [vp, C.POINTER(SketchCsr)]),
    "suminc_pr_step_phase": (i32, [vp, i32, C.POINTER(C.c_double)]),
    "suminc_pr_correct": (i32, [vp]),
    "suminc_pr_precompute": (i32, [vp]),
    "suminc_traversal_init": (i32, [vp, C.c_int, u32]),
    "suminc_cluster_lpa": (i32, [u32, vp, vp, u32, u32, u32, u32, vp, vp]),
    "suminc_compress_mirror_save": (i32, [vp, C.c_char_p]),
    "suminc_compress_mirror_load": (i32, [vp, C.c_char_p, vp]),
    "suminc_serialization_path": (i32, [C.c_char_p, C.c_char_p, C.c_char_p, u32, u32, C.c_char_p, u64, vp]),
    "suminc_edges_save": (i32, [vp, C.c_char_p]),
    "suminc_edges_load": (i32, [C.c_char_p, vp]),
    "suminc_updates_save": (i32, [C.c_char_p, vp, u64]),
    "suminc_updates_load": (i32, [C.c_char_p, vp, u64, vp]),
    "suminc_php_run": (i32, [vp, u32, C.c_double, vp, vp]),
    "suminc_sketch_pr_run": (i32, [vp, C.POINTER(Stats)]),
    "suminc_compress_mirror": (i32, [vp, vp, u32, u32, C.c_double, C.c_double, C.POINTER(CompressInfo)]),
    "suminc_sketch_download": (i32, [vp, C.POINTER(SketchBuffers)]),
    "suminc_get_stats": (i32, [vp, C.POINTER(Stats)]),
    "suminc_cluster_leaves": (i32, [vp, u32, u32, vp]),
    "suminc_set_clusters": (i32, [vp, vp]),
    "suminc_sum_build": (i32, [vp, u32, u32, f64]),
    "suminc_sum_info": (i32, [vp, C.POINTER(SketchInfo)]),
    "suminc_sum_download": (i32, [vp, vp, vp]),
    "suminc_sum_pr_shortcut": (i32, [vp, u32, u32, vp, vp, vp, C.POINTER(u32), C.POINTER(u32)]),
    "suminc_sum_pr_run": (i32, [vp, C.POINTER(Stats)]),
    "suminc_sum_sssp_shortcut": (i32, [vp, u32, u32, vp, vp, vp, C.POINTER(u32), C.POINTER(u32)]),
    "suminc_sum_sssp_run": (i32, [vp, C.POINTER(Stats)]),
    "suminc_wpr_set_weights": (i32, [vp, vp]),
    "suminc_wpr_forward": (i32, [vp, vp, vp, C.POINTER(Stats)]),
    "suminc_wpr_forward_noalloc": (i32, [vp, vp, vp, C.POINTER(Stats)]),
    "suminc_wpr_backward": (i32, [vp, vp, vp, vp, C.POINTER(Stats)]),
    "suminc_wpr_download": (i32, [vp, vp, vp]),
    "suminc_profile_enable": (i32, [vp, i32]),
    "suminc_profile_get": (i32, [vp, i32, C.POINTER(KernelProfile)]),
    "suminc_edges_read_tsv": (i32, [C.c_char_p, i32, i32, C.POINTER(vp)]),
    "suminc_edges_info": (i32, [vp, C.POINTER(u32), C.POINTER(u64), C.POINTER(i32), C.POINTER(u64)]),
    "suminc_edges_copy": (i32, [vp, vp, vp, vp]),
    "suminc_edges_upload": (i32, [vp, vp]),
    "suminc_edges_free": (None, [vp]),
    "suminc_updates_read_tsv": (i32, [C.c_char_p, i32, i32, vp, u64, C.POINTER(u64)]),
    "suminc_clusters_read": (i32, [C.c_char_p, u32, u32, vp, C.POINTER(u32)]),
    "suminc_clusters_write": (i32, [C.c_char_p, u32, vp]),
}

KERNEL_PR_PUSH = 0
KERNEL_PR_GATHER = 1
KERNEL_SSSP_PUSH = 2
KERNEL_PR_PULL = 3
KERNEL_SSSP_SKETCH = 4
KERNEL_BATCH_COPY = 5


def lib():
    """Loads libsuminc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `python suminc_amd/build.py` (the device path has no CPU fallback)"
            )
        L = C.CDLL(str(LIB_PATH), mode=getattr(os, "RTLD_NOW", 2))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to libsuminc must be C-contiguous"
    return a.ctypes.data_as(vp)
