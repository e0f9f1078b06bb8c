import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle():
    from helpers import Oracle
    return Oracle()


@pytest.fixture
def odirect_state_dir(tmp_path):
    """(directory accepting O_DIRECT or None, refusals): see helpers.odirect_dir.
    A folder made outside tmp_path is removed after the test."""
    import shutil
    from helpers import odirect_dir
    d, seen = odirect_dir(tmp_path)
    yield d, seen
    if d is not None and d != str(tmp_path):
        shutil.rmtree(d, ignore_errors=True)


def _gpu_present():
    # The kernel driver node; checking it does not initialise the GPU.
    return os.path.exists("/dev/kfd")


def pytest_collection_modifyitems(config, items):
    if _gpu_present():
        return
    skip = pytest.mark.skip(reason="no GPU in this container")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
