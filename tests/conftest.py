import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpicugpu.so)")
    config.addinivalue_line("markers", "slow: long CPU-only oracle runs")


@pytest.fixture(scope="session")
def gpu():
    """The GPU engine, initialised once per session (fails loudly if absent)."""
    from piculator_amd import picugpu

    picugpu.init(0)
    return picugpu
