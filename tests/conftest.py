import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import numpy as np

DAMPING = float(np.float32(0.85))  # (double)0.85f, PageRankIngress::d


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


def has_gpu_node() -> bool:
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu():
    if not has_gpu_node():
        pytest.skip("no GPU device node (/dev/kfd) in this environment")
    import suminc_amd

    return suminc_amd


def pytest_collection_modifyitems(config, items):
    # print the case name before each GPU test so a crash names it
    for item in items:
        if "gpu" in item.keywords:
            item.user_properties.append(("gpu", True))


@pytest.fixture(autouse=True)
def _announce(request):
    if "gpu" in request.node.keywords:
        print(f"\n[gpu-case] {request.node.nodeid}", flush=True)
    yield
