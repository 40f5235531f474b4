import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a HIP device (MI355X); runs through libaxs_hip.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle_py
    oracle_py.load()
    return oracle_py


@pytest.fixture(scope="session")
def ref():
    """The reference's own standalone headers compiled from source, or skip."""
    import oracle_py
    lib = oracle_py.ref_lib()
    if lib is None:
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    return lib


@pytest.fixture(scope="session")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from axion_strings_adaptive_amd import capi
    c = capi.Context(0)
    yield c
    c.close()
