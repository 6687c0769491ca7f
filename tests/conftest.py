import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "hermite-ugks_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def have_gpu():
    from hugks import capi
    lib = capi.load()
    info = lib.hugks_device_info().decode()
    if info.startswith("no HIP device"):
        pytest.skip("no HIP device")
    return info
