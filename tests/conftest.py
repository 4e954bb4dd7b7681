import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU oracle check")


@pytest.fixture(scope="session")
def ctx64():
    from polystring_amd.capi import Context
    # objects created from a context keep a reference to it; the context is
    # released when the last of them is collected
    return Context(0, "fp64")


@pytest.fixture(scope="session")
def ctx32():
    from polystring_amd.capi import Context
    return Context(0, "fp32")
