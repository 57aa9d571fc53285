import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "hr-weno_amd")
for p in (PKG, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on the device)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    if not os.path.exists(o.ORACLE_SO):
        o.build(ref=False)
    return o.Oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle as o
    if not o.Reference.available():
        pytest.skip("oracle/_ref (the reference's own build) is not present")
    return o.Reference()


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def hh():
    import hrweno_hip
    return hrweno_hip
