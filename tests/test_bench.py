"""bench.py's command-line contract, checked without a GPU: more than one GPU is
refused with a message, every workload names its kind and sizes, and the
per-workload traffic summaries the roofline reads are found."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120, cwd=str(ROOT))
    assert out.returncode == 2
    assert "one GPU" in out.stderr
    assert out.stdout == ""


def test_workloads_cover_the_configurations():
    kinds = {w["kind"] for w in bench.WORKLOADS.values()}
    assert kinds == {"sssp", "pr", "prfull", "wpr", "web"}
    for name, w in bench.WORKLOADS.items():
        assert w["scale"] in (16, 22, 24, 26), name
        assert w["desc"], name
    assert bench.WORKLOADS["sssp22"]["frac"] == 0.001 and bench.WORKLOADS["sssp22"]["max_node"] == 1024
    # the default workload is the metric's large configuration: s26, seed 5, 1% batches,
    # supernodes, SSSP weights U[1,255]
    w = bench.WORKLOADS["web26"]
    assert (w["scale"], w["seed"], w["frac"], w["max_node"], w["wmax"]) == (26, 5, 0.01, 1024, 255)
    src = (ROOT / "bench.py").read_text()
    assert 'ap.add_argument("--workload", default="web26"' in src


def test_traffic_summaries_are_found():
    assert bench.measured_traffic("k_copy_old", "sssp22") > 1e9
    assert bench.measured_traffic("k_pr_pull", "pr22") > 1e9
    assert bench.measured_traffic("k_pr_pull", "pr26") > 1e10
    assert bench.measured_traffic("k_pr_pull", "pr16") is None


def test_pull_mode_follows_the_library_rules():
    """The kernel the roofline names is the one the library's dense rounds run:
    binned above 8 MiB of FP32 contributions (pr.hip build_bins), else tiled above
    64 MiB (build_tiles), else the row classes; a window given forces its pull,
    a negative one turns it off, and the banded pull keeps the row classes."""
    from types import SimpleNamespace

    def mode(n, band=0, tile=0, binned=0):
        return bench.pull_mode(SimpleNamespace(pull_band_kib=band, pull_tile_kib=tile, pull_bin_kib=binned), n)

    assert mode(1 << 16) == "rows"
    assert mode(1 << 21) == "rows"  # exactly 8 MiB: not above the threshold
    assert mode((1 << 21) + 1) == "binned"
    assert mode(1 << 22) == "binned" and mode(1 << 26) == "binned"
    assert mode(1 << 26, binned=-1) == "tiled"
    assert mode(1 << 24, binned=-1) == "rows"  # 64 MiB: not above the tiled threshold
    assert mode(1 << 16, binned=4) == "binned" and mode(1 << 16, binned=-1, tile=4) == "tiled"
    assert mode(1 << 26, band=32768) == "rows"
    assert mode(1 << 26, binned=-1, tile=-1) == "rows"
    assert bench.PULL_KERNELS["binned"] == "k_pr_bin_gather+k_pr_bin_sum"
