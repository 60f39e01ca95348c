import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build(ref=False) if not os.path.exists(os.path.join(ROOT, "oracle", "libicon_oracle.so")) else None
    return oracle


@pytest.fixture(scope="session")
def icon():
    """The HIP library; GPU tests fail (not skip) when it is missing."""
    import feilab_amd
    feilab_amd.lib()
    return feilab_amd
