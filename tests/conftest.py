import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libofx_hip.so")


@pytest.fixture(scope="session")
def emu_ctx():
    """Context over the CPU emulation of the kernel (tests/emu/libofx_emu.so)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "emu")])
    from optifax_amd.solver import Context
    ctx = Context(lib_path=os.path.join(ROOT, "tests", "emu", "libofx_emu.so"))
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def gpu_ctx():
    from optifax_amd.solver import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
