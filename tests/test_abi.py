"""The C ABI, on the CPU: the HIP library loads and exports every function include/lmw.h declares (no compute
calls without a GPU), and the Python mirror binds exactly those names."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names += re.findall(r"^(?:lmw_status|const char \*)\s*(lmw_\w+)\s*\(", text, flags=re.M)
    return names


def test_header_declares_the_surface():
    names = declared()
    assert len(names) == len(set(names))
    for required in ("lmw_sphere_create", "lmw_sphere_run", "lmw_sphere_step", "lmw_plane_run", "lmw_ensemble_steps_graph"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from lmweather_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("liblmw_hip.so is not built: run python -m lmweather_amd.build")
    so = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(so, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from lmweather_amd import _lib
    names = set(declared()) - {"lmw_last_error"}
    assert names <= set(_lib.SIGNATURES), sorted(names - set(_lib.SIGNATURES))


def test_product_does_not_touch_the_oracle():
    """The product package never imports, loads or runs anything under oracle/."""
    pkg = os.path.join(ROOT, "lmweather_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text and "liblmw_oracle" not in text, f
