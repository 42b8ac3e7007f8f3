"""Pinning of the oracle's expression compiler (oracle/expr.py), on the CPU:
  * against the reference's own compiler (cpp/expr.hpp, compiled into oracle/_ref): values at random points and the
    exact error texts of malformed strings, including the cases of test.sh:179-215;
  * against Python's evaluation of the same strings (Python's grammar for signs, +, -, *, / and ** is the
    reference's), on seeded random expressions: precedence and associativity checked without either compiler, bit
    for bit (both sides run numpy's array loops);
  * closed-form values.
"""
import math
import random

import numpy as np
import pytest

from oracle import expr as X

@pytest.fixture(scope="module")
def ref():
    if not X.ref_available():
        pytest.skip("oracle/_ref not built (the reference tree is absent)")
    return X.ref_eval


SPHERE_PAR = X.sphere_params(6.37122e6, 7.292e-5, 9.80616)
PLANE_PAR = X.plane_params(1.0e6, 5.0e5, 100.0, 9.81, 1e-4, 1.6e-11)
CONE = "2000.0 * (1 - min(pi / 9, sqrt((lon - 1.5 * pi) * (lon - 1.5 * pi) + (lat - pi / 6) * (lat - pi / 6))) / (pi / 9))"
SPHERE_EXPRS = [
    "1000", "2000.0", "15.0 * cos(lat)", "6.0 * cos(3.0 * lon) * cos(lat) * cos(lat)",
    "5960.0 - 100.0 * pow(2.718281828459045, -((lon * cos(lat)) * (lon * cos(lat)) + (lat - 0.7853981633974483) * (lat - 0.7853981633974483)) / 0.04)",
    "29400.0 / g - (a * omega * 38.6106 + 38.6106 * 38.6106 / 2.0) * pow(0 - cos(lon) * cos(lat) * sin(1.5) + sin(lat) * cos(1.5), 2) / g",
    "38.6106 * (cos(lat) * cos(1.5) + sin(lat) * cos(lon) * sin(1.5))", "0 - 38.6106 * sin(lon) * sin(1.5)",
    "500.0 * (1 + cos(pi * min(1.0, 3.0 * acos(cos(lat) * cos(lon - 1.5 * pi)))))", "1.0 + sin(lat)", CONE,
    "5960.0 - " + CONE,
    "tan(lat) + atan(lon) + sinh(lat) + cosh(lat) + tanh(lon) + exp(lat) + log(lon + 1) + ln(2 + lat) + log10(lon + 3)",
    "asin(lat / 2) + abs(lat) + floor(3 * lon) + ceil(2 * lat) + sign(lat) + atan2(lat, lon) + max(lat, lon) + mod(7.5 * lon, 1.3)",
    "-lat^2 + 2^-lon + lon**2**0.5 - -lat - +lon * 1e-3 / .5e1", "(((lon)))", "1e300 * 1e300", "0 / (lat - lat)",
    "sqrt(lat)", "log(lat - lat)", "0x10 + lon", "  1 +\t2 \n",
]
PLANE_EXPRS = [
    "exp(-(y - ly / 2) * (y - ly / 2) / 160000000000.0)", "cos(2 * pi * x / lx)", "0.5 * 31.3209195267 / 100.0 * cos(2 * pi * x / lx)",
    "4.0 * sin(4 * pi * x / lx) * sin(pi * y / ly)", "f0 * 100000.0 * cos(2 * pi * x / lx) * sin(pi * y / ly) / g",
    "-100000.0 * pi / ly * cos(2 * pi * x / lx) * cos(pi * y / ly)", "H + beta * y", "-150.0 * cos(2.0 * pi * x / lx)",
]


def close(a, b):
    """Same non-finite pattern, finite values within 8 ulp of the larger magnitude in the expression."""
    a, b = np.asarray(a), np.asarray(b)
    if not np.array_equal(np.isnan(a), np.isnan(b)) or not np.array_equal(np.isinf(a), np.isinf(b)):
        return False
    f = np.isfinite(b)
    if np.any(np.isinf(b) & (np.sign(a) != np.sign(b))):
        return False
    scale = np.maximum(np.abs(b[f]), 1e-300)
    return bool(np.all(np.abs(a[f] - b[f]) <= 8 * np.spacing(scale) + 1e-15 * np.max(np.abs(b[f]), initial=0.0)))


def test_values_match_reference_compiler(ref):
    rng = np.random.default_rng(1)
    lon = rng.uniform(0, 2 * np.pi, 300)
    lat = rng.uniform(-1.5, 1.5, 300)
    for t in SPHERE_EXPRS:
        want = ref(t, ["lon", "lat"], SPHERE_PAR, lon, lat)
        got = X.Expr(t, ["lon", "lat"], SPHERE_PAR).eval(lon, lat)
        assert close(np.broadcast_to(got, want.shape), want), t
    x = rng.uniform(0, 1e6, 300)
    y = rng.uniform(0, 5e5, 300)
    for t in PLANE_EXPRS:
        want = ref(t, ["x", "y"], PLANE_PAR, x, y)
        got = X.Expr(t, ["x", "y"], PLANE_PAR).eval(x, y)
        assert close(np.broadcast_to(got, want.shape), want), t


BAD = [
    "2 +* 3", "foo", "foo(1)", "sin(1, 2)", "atan2(1)", "pow()", "(1 + 2", "1 +", "", "   ", ".", "..5", "1 2", "2 ** * 3",
    "@", "1e400", "1e-400", "sin", "lat)", "min(1,)", "mod(1 2)", "3 $ 4", "x_y + 1", "_a", "1 + 2 + " + "3 * " * 40 + "foo",
    "(" * 300 + "1" + ")" * 300, "(" * 99 + "1" + ")" * 99 + ")", "-" * 199 + "1", "-" * 200 + "1",
    "+".join(["1"] * 2000), "+".join(["1"] * 2001), "1 +\n\t@",
]


def test_error_texts_match_reference_compiler(ref):
    one = np.zeros(1)
    for t in BAD:
        try:
            ref(t, ["lon", "lat"], SPHERE_PAR, one, one)
            want = None
        except X.ExprError as e:
            want = str(e)
        try:
            X.Expr(t, ["lon", "lat"], SPHERE_PAR)
            got = None
        except X.ExprError as e:
            got = str(e)
        assert got == want, t


def test_reference_error_cases():
    """test.sh:179-180,204-206: the messages the reference's script tests look for."""
    for t, part in [("2 +* 3", "expression"), ("foo", "unknown name"), ("(" * 300 + "1" + ")" * 300, "nested too deeply"),
                    ("1e400", "out of range")]:
        with pytest.raises(X.ExprError, match=part):
            X.Expr(t, ["lon", "lat"], SPHERE_PAR)
    with pytest.raises(X.ExprError) as e:
        X.Expr("foo", ["lon", "lat"], SPHERE_PAR)
    assert str(e.value) == "expression \"foo\": unknown name 'foo' at column 4"


def test_closed_forms():
    def ev(t, lon=0.0, lat=0.0):
        return float(X.Expr(t, ["lon", "lat"], SPHERE_PAR).eval(lon, lat))
    assert ev("-2^2") == -4.0 and ev("2^3^2") == 512.0 and ev("2**-1") == 0.5 and ev("(-2)^2") == 4.0
    assert ev("10/4/5") == 0.5 and ev("8-3-2") == 3.0 and ev("2*3+4*5") == 26.0 and ev("2+3*4^2") == 50.0
    assert ev("sign(-3)") == -1.0 and ev("sign(0)") == 0.0 and ev("min(2, 1)") == 1.0 and ev("max(2, 1)") == 2.0
    assert ev("mod(-7, 3)") == -1.0 and ev("abs(-2.5)") == 2.5 and ev("floor(-1.5)") == -2.0 and ev("ceil(-1.5)") == -1.0
    assert ev("atan2(1, -1)") == math.atan2(1, -1) and ev("pi") == math.pi and ev("e") == math.e
    assert ev("a") == 6.37122e6 and ev("g") == 9.80616 and ev("omega") == 7.292e-5
    assert ev("lon * 10 + lat", 2.0, 3.0) == 23.0 and ev("0x10") == 16.0 and ev(".5e1") == 5.0
    assert math.isnan(ev("sqrt(-1)")) and ev("1 / 0") == math.inf and ev("-1 / 0") == -math.inf


def _random_expr(rng, depth):
    """(reference syntax, Python syntax) of a random well-formed expression over lon and lat."""
    def atom(d):
        r = rng.random()
        if d <= 0 or r < 0.35:
            if rng.random() < 0.5:
                x = round(rng.uniform(0.1, 3.0), 3)
                return f"{x}", f"F({x})"
            n = rng.choice(["lon", "lat"])
            return n, n
        if r < 0.6:
            a, b = expr(d - 1)
            return f"({a})", f"({b})"
        f = rng.choice(["sin", "cos", "exp", "atan", "sqrt", "abs", "min", "max", "atan2"])
        if f in ("min", "max", "atan2"):
            a1, b1 = expr(d - 1)
            a2, b2 = expr(d - 1)
            return f"{f}({a1}, {a2})", f"{f}({b1}, {b2})"
        a, b = expr(d - 1)
        return f"{f}({a})", f"{f}({b})"

    def signed(d):
        r = rng.random()
        if r < 0.15:
            a, b = signed(d)
            return f"-{a}", f"-{b}"
        if r < 0.2:
            a, b = signed(d)
            return f"+ {a}", f"+ {b}"
        return power(d)

    def power(d):
        a, b = atom(d)
        if rng.random() < 0.2:
            c, e = signed(d - 1)
            op = rng.choice(["^", "**"])
            return f"{a} {op} {c}", f"{b} ** {e}"
        return a, b

    def product(d):
        a, b = signed(d)
        for _ in range(rng.randrange(3)):
            op = rng.choice("*/")
            c, e = signed(d)
            a, b = f"{a} {op} {c}", f"{b} {op} {e}"
        return a, b

    def expr(d):
        a, b = product(d)
        for _ in range(rng.randrange(3)):
            op = rng.choice("+-")
            c, e = product(d)
            a, b = f"{a} {op} {c}", f"{b} {op} {e}"
        return a, b

    return expr(depth)


def test_grammar_against_python():
    rng = random.Random(7)
    ns = {"F": lambda x: np.array([x]), "sin": np.sin, "cos": np.cos, "exp": np.exp, "atan": np.arctan, "sqrt": np.sqrt, "abs": np.fabs,
          "min": lambda a, b: np.where(a < b, a, b), "max": lambda a, b: np.where(a > b, a, b), "atan2": np.arctan2}
    pts = [(np.array([0.7]), np.array([-0.4])), (np.array([2.5]), np.array([1.1]))]  # both sides on numpy's array loops
    for _ in range(400):
        ref_text, py_text = _random_expr(rng, 3)
        e = X.Expr(ref_text, ["lon", "lat"], SPHERE_PAR)
        for lon, lat in pts:
            with np.errstate(all="ignore"):
                want = np.broadcast_to(eval(py_text, dict(ns, lon=lon, lat=lat)), (1,))
            got = np.broadcast_to(e.eval(lon, lat), (1,))
            assert np.array_equal(got, want, equal_nan=True), (ref_text, got, want)
