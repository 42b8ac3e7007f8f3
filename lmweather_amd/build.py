"""Build recipe of the gfx950 library: compiles lmweather_amd/csrc/*.hip with hipcc into lmweather_amd/liblmw_hip.so.

Run as ``python -m lmweather_amd.build`` (or through ``__graft_entry__.build()``).  The library is built in-tree so it
travels with the repository; it exports exactly the C ABI declared in include/lmw.h.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblmw_hip.so")
ARCH = os.environ.get("LMW_OFFLOAD_ARCH", "gfx950")

SOURCES = ["sphere_core.hip", "sphere_abi.hip", "plane.hip", "plane_f32.hip", "precision.hip", "corrector.hip", "autotune.hip", "output.hip", "expr.hip"]

FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden",
    "-Wall",
    "-Wno-unused-function",
]


def sources() -> list[str]:
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compile every source to an object in parallel, then link the shared library.  Returns its path."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    procs, objs = [], []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [hipcc, *FLAGS, *inc, "-DLMW_BUILDING", "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    tmp = LIB + ".tmp"
    cmd = [hipcc, *FLAGS, "-shared", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}")
    os.replace(tmp, LIB)
    return LIB


CPP_TESTS = ("facade_checks", "ops_checks")


def build_cpp_tests() -> list[str]:
    """Compile the native test programs into tests/bin, linked against liblmw_hip.so: tests/facade_checks.cpp
    (closed-form checks of the C++ facade include/lmw.hpp) and tests/ops_checks.cpp (the operator table
    include/lmw_ops.hpp: schema, signatures, every operator once)."""
    out_dir = os.path.join(ROOT, "tests", "bin")
    os.makedirs(out_dir, exist_ok=True)
    built = []
    for name in CPP_TESTS:
        exe = os.path.join(out_dir, name)
        cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
               os.path.join(ROOT, "tests", name + ".cpp"), "-L", HERE, "-llmw_hip", "-Wl,-rpath,$ORIGIN/../../lmweather_amd", "-o", exe]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"building {exe} failed:\n{r.stdout}")
        built.append(exe)
    return built


def _drain(procs) -> None:
    errors = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            errors.append(f"{os.path.basename(src)}:\n{out}")
        elif out.strip():
            sys.stderr.write(out)
    procs.clear()
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_cpp_tests())
