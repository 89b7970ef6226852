"""suminc_amd -- MI355X (gfx950) implementation of SumInc's hot path.

Delta-accumulative PageRank and SSSP over a CSR fragment in HBM, optionally
compressed into supernodes with entry/exit shortcuts, applied incrementally
after edge-update batches.  The device code and the C ABI live in csrc/ and are
built into libsuminc.so (include/suminc.h); this package is the Python handle.
"""
from ._lib import (  # noqa: F401
    APP_BFS,
    APP_SSSP,
    APP_SSWP,
    E_INVALID,
    E_NOMEM,
    E_NOT_CONVERGED,
    E_NUMERIC,
    E_STATE,
    E_TOO_LARGE,
    INF_U32,
    OP_ADD,
    OP_DEL,
    PHASE_PLAIN,
    PHASE_SKETCH,
    PR_FP32_KAHAN,
    PR_FP64,
    UPDATE_DTYPE,
    SumincError,
)
from .context import Context  # noqa: F401

__all__ = ["Context", "SumincError", "UPDATE_DTYPE", "INF_U32", "OP_ADD", "OP_DEL", "PR_FP64", "PR_FP32_KAHAN",
           "E_INVALID", "E_NOMEM", "E_NOT_CONVERGED", "E_NUMERIC", "E_STATE", "E_TOO_LARGE", "APP_SSSP", "APP_BFS",
           "APP_SSWP", "PHASE_PLAIN", "PHASE_SKETCH"]
