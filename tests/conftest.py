import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libbs_voxel.so")


@pytest.fixture(scope="session")
def oracle_lib():
    """The CPU oracle, built once if it is not there yet."""
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    O.lib()
    return O


@pytest.fixture(scope="session")
def gpu_ctx():
    """A bs_ctx on device 0; the HIP library must be built and load."""
    import baby_shark_amd as bs
    bs.load_library()
    return bs.default_context(0)
