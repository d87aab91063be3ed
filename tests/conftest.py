import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs through librustracer_hip.so)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def rt():
    import rustracer_amd

    rustracer_amd.lib()
    return rustracer_amd


def pytest_runtest_makereport(item, call):
    """Stop the session after a HIP runtime error: nothing more runs on a GPU
    that may have faulted."""
    if call.excinfo is not None:
        err = call.excinfo.value
        if getattr(err, "code", None) == -3:  # RT_ERR_HIP
            item.session.shouldstop = f"HIP runtime error in {item.nodeid}: {err}"
