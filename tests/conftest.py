"""Shared fixtures.  Tests marked ``gpu`` need an MI355X and the built liblmw_hip.so; the rest run on the CPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and the built HIP library")


def _ensure_oracle():
    so = os.path.join(ROOT, "oracle", "liblmw_oracle.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liblmw_oracle.so"], check=True)


_ensure_oracle()


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyapi
    return pyapi

