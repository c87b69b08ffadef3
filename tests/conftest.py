"""Test configuration: the `gpu` marker and import paths.

`-m "not gpu"` runs the CPU suite (oracle vs golden vectors / closed forms, host
logic, C-ABI exports); `-m gpu` runs the HIP parity tests through the C-ABI.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "vf-version_7_3_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def orc():
    import orc as _orc
    _orc.lib()
    return _orc


@pytest.fixture(scope="session")
def vfd():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import vfd as _vfd
    return _vfd
