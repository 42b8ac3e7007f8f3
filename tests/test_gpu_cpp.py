"""The reference's native core tests (reference tests/core.cpp) restated in C++ on the façade include/lmw.hpp and run
against the GPU library: the C++ host path of the drop-in."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin", "core_gpu")


def test_core_checks_on_the_cpp_facade():
    if not os.path.exists(EXE):
        pytest.fail("tests/bin/core_gpu is not built: run __graft_entry__.build()")
    r = subprocess.run([EXE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "core tests passed" in r.stdout
