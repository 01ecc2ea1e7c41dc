import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "freestream-milne_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (ROOT, PKG, ORACLE, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs through libfsm.so)")


def _ensure_built():
    # Load what is built; build only when nothing is.
    if not os.path.exists(os.path.join(PKG, "libfsm.so")):
        subprocess.check_call(["make", "-C", PKG, "-s", "libfsm.so"])
    if not os.path.exists(os.path.join(ORACLE, "_build", "liboracle.so")):
        subprocess.check_call(["make", "-C", ORACLE, "-s"])


_ensure_built()


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch  # noqa: F401  (one HIP runtime for torch and libfsm)
    import fsm_amd

    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    ctx = fsm_amd.Context(0)
    yield ctx
    ctx.close()
