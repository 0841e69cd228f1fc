"""Shared test setup.

`gpu` tests need a gfx950 device and call libyune_hip through its C ABI;
everything else runs on the CPU (oracle, host-side C++, ABI surface).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
REF_ROOT = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and runs the HIP path")


@pytest.fixture(scope="session")
def ctx():
    import yune_amd as Y
    c = Y.Context()
    yield c
    c.close()
