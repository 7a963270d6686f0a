import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE = os.path.join(ROOT, "oracle")
if ORACLE not in sys.path:
    sys.path.insert(0, ORACLE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU run")


@pytest.fixture(autouse=True)
def _no_planted_fault():
    """Every test starts and ends with the oracle's fault hook cleared and its
    reductions correctly rounded."""
    import oracle_ffi
    oracle_ffi.lib().orc_set_fault(0)
    oracle_ffi.lib().orc_set_sum_mode(0)
    yield
    oracle_ffi.lib().orc_set_fault(0)
    oracle_ffi.lib().orc_set_sum_mode(0)
