"""The C ABI library loads and exports every function include/suminc.h declares;
argument/state errors come back as codes.  CPU only (no compute calls)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from suminc_amd import _lib as L

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "suminc.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(suminc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (suminc_[a-z0-9_]+)", out))
    assert set(names) <= exported
    # every declared function has a ctypes signature in the binding
    assert set(names) <= set(L.SIGNATURES)


def test_default_params_follow_reference_flags():
    p = L.Params()
    L.lib().suminc_default_params(C.byref(p))
    assert p.damping == pytest.approx(0.8500000238418579, abs=0)  # (double)0.85f
    assert p.pr_precision == L.PR_FP32_KAHAN
    assert p.max_rounds > 0


def test_null_arguments_return_codes():
    lib = L.lib()
    assert lib.suminc_ctx_create(0, None, None) == L.E_INVALID
    assert lib.suminc_ctx_destroy(None) == L.E_INVALID
    assert lib.suminc_pr_init(None) == L.E_INVALID
    # without a context: the last failed context-free call of this thread (none yet here)
    assert lib.suminc_last_error(None) is not None


def test_context_creation_without_gpu_fails_cleanly():
    if Path("/dev/kfd").exists():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = L.lib().suminc_ctx_create(0, None, C.byref(h))
    assert rc in (L.E_HIP, L.E_INVALID)
    assert not h.value


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors have the C structs' sizes and field offsets (a C probe
    compiled against include/suminc.h prints them)."""
    import numpy as np

    mirrors = {"suminc_params": L.Params, "suminc_stats": L.Stats, "suminc_sketch_info": L.SketchInfo,
               "suminc_kernel_profile": L.KernelProfile, "suminc_sketch_csr": L.SketchCsr,
               "suminc_compress_info": L.CompressInfo, "suminc_sketch_buffers": L.SketchBuffers}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "suminc.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append('  printf("suminc_update size %zu\\n", sizeof(suminc_update));')
    for f in L.UPDATE_DTYPE.names:
        lines.append(f'  printf("suminc_update {f} %zu\\n", offsetof(suminc_update, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        s, f, v = line.split()
        got[(s, f)] = int(v)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
    dt = np.dtype(L.UPDATE_DTYPE)
    assert got[("suminc_update", "size")] == dt.itemsize
    for f in dt.names:
        assert got[("suminc_update", f)] == dt.fields[f][1], f
