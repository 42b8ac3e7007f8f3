"""The C++ facade include/lmw.hpp (the reference's Sphere/Plane class API over the C ABI) driven by a native program,
tests/facade_checks.cpp: closed forms and exact properties with the bounds of the reference's own native test
(reference tests/core.cpp:17-121) -- the C++ host path of the drop-in."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin", "facade_checks")


def test_facade_checks():
    if not os.path.exists(EXE):
        pytest.fail("tests/bin/facade_checks is not built: run __graft_entry__.build()")
    r = subprocess.run([EXE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "facade checks passed" in r.stdout
