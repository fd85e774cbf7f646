import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (libnumfast_hip kernels run)")
    config.addinivalue_line("markers", "slow: larger sizes, still bounded")


@pytest.fixture(scope="session")
def c_oracle():
    from oracle.executor import load_c_oracle
    lib, groups = load_c_oracle()
    return lib


@pytest.fixture(scope="session")
def hip():
    """The product library, loaded through the same binding as NumFast's
    loader.  Only GPU tests request this fixture."""
    os.environ.pop("NUMFAST_NATIVE_DISABLE", None)
    # torch first: loading the library before torch initialises the HIP
    # runtime in a way that leaves torch without a visible GPU
    import torch  # noqa: F401
    from numfast_amd import native
    lib = native.lib()
    if not lib.nfh_device_ok():
        pytest.fail("libnumfast_hip loaded but no usable GPU")
    return lib
