"""The operator table include/lmw_ops.hpp, on the CPU (no GPU call is made): the schema document it generates is
byte for byte the reference's (sha256 of reference schema/lmweather.dsl.json, which script bytecode binds to:
reference cpp/dslvm.hpp:31), and its canonical signature strings are those the reference schema declares (81 host
methods: Wx 8, Sphere 44, Plane 29, and the context wx: Wx).  Fixture: tests/golden/schema_fixture.json, written by
tests/golden/make_schema_fixture.py from the reference's schema file."""
import hashlib
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "bin", "ops_checks")
FIXTURE = json.load(open(os.path.join(ROOT, "tests", "golden", "schema_fixture.json")))


def run(arg):
    if not os.path.exists(EXE):
        pytest.fail("tests/bin/ops_checks is not built: run __graft_entry__.build()")
    r = subprocess.run([EXE, arg], stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=60)
    assert r.returncode == 0, r.stderr.decode()
    return r.stdout


def test_schema_is_byte_identical():
    schema = run("--schema")
    assert len(schema) == FIXTURE["bytes"]
    assert hashlib.sha256(schema).hexdigest() == FIXTURE["sha256"]
    doc = json.loads(schema)
    assert {t: len(b["methods"]) for t, b in doc["types"].items()} == {"Wx": 8, "Sphere": 44, "Plane": 29}


def test_signatures_match_the_reference_schema():
    lines = run("--signatures").decode().splitlines()
    assert lines == FIXTURE["signatures"]
    assert sum(1 for s in lines if s.startswith("method ")) == 81
