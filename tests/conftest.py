import glob
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libiccg_hip.so")


def gpu_visible():
    """A GPU node is present.  Decided without loading the HIP library, so a
    missing or broken library fails the GPU tests instead of skipping them."""
    return os.path.exists("/dev/kfd") and bool(glob.glob("/dev/dri/renderD*"))


def pytest_collection_modifyitems(config, items):
    if gpu_visible():
        return
    skip = pytest.mark.skip(reason="no GPU device node (/dev/kfd) visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
