"""Every Sphere and Plane operator of the table include/lmw_ops.hpp called once through the table on the GPU
(tests/ops_checks.cpp): results against the façade called directly and closed forms, the registration wrapper's
refusals (argument kinds, non-finite results, the reference's argument checks and texts, reference cpp/main.cpp:62-80,
108-495)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin", "ops_checks")


def test_every_operator_through_the_table(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/bin/ops_checks is not built: run __graft_entry__.build()")
    r = subprocess.run([EXE, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "operator checks passed (81 operators)" in r.stdout
    for f in ("sphere.csv", "spectrum.txt", "zonal.txt", "plane.csv", "row.txt", "column.txt", "text.txt"):
        assert (tmp_path / f).stat().st_size > 0, f
