import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "idrs-package_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a visible GPU; fails (does not skip) without one."""
    import idrs_amd

    lib = idrs_amd.load_library()
    n = lib.idrs_device_count()
    assert n >= 1, "no HIP device visible: GPU tests need an MI355X"
    return lib
