"""A C program (tests/c/worker_phases.c) drives the library through the reference
worker's phase order with nothing but the C ABI: it compiles and links here, and
runs on the GPU."""
import subprocess
from pathlib import Path

import pytest

from suminc_amd import _lib as L

ROOT = Path(__file__).resolve().parent.parent


def _build(tmp_path):
    exe = tmp_path / "worker_phases"
    libdir = str(Path(L.LIB_PATH).resolve().parent)
    subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-Werror", "-I", str(ROOT / "include"),
                    str(ROOT / "tests" / "c" / "worker_phases.c"), "-o", str(exe), "-L", libdir, "-lsuminc",
                    f"-Wl,-rpath,{libdir}", "-lm"], check=True)
    return exe


def test_worker_driver_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run(["nm", "-u", str(exe)], capture_output=True, text=True, check=True).stdout
    for name in ("suminc_pr_precompute", "suminc_pr_step_phase", "suminc_pr_correct", "suminc_upload_sketch",
                 "suminc_compress_mirror", "suminc_sketch_download", "suminc_sssp_step"):
        assert name in out


@pytest.mark.gpu
def test_worker_driver_runs(gpu, tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().endswith("worker_phases: PASS")
