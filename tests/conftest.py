import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle():
    import oracle as _oracle

    _oracle.lib()
    return _oracle


@pytest.fixture(scope="session")
def ref_lib():
    import oracle as _oracle

    if not _oracle.ref_available():
        pytest.skip("reference CPU build (oracle/_ref) not built on this host")
    return _oracle
