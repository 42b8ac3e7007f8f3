"""TEST INFRASTRUCTURE ONLY: a CPU restatement of the reference's expression compiler (cpp/expr.hpp:16-166) and of
the grids its operators evaluate (Sphere.fields / orography / tracer, cpp/main.cpp:158-172,212-228,252-268;
Plane.fields, cpp/main.cpp:410-431), used by the tests as the checker of the library's device evaluation.

The grammar, the names, the functions, the limits and the error texts are the reference's:
  sum     := product (('+' | '-') product)*            (cpp/expr.hpp:64-72, nesting counted)
  product := signed (('*' | '/') signed)*              (:73-80)
  signed  := ('-' | '+') signed | power                (:81-86, nesting counted)
  power   := atom (('^' | '**') signed)?               (:87-94, right associative)
  atom    := '(' sum ')' | number | name '(' args ')' | name   (:95-129, leaves counted)
Numbers are read with the C library's strtod, as std::stod does (cpp/expr.hpp:103-105).  Evaluation is vectorised
over numpy arrays: one IEEE operation per node, the C library's pow and fmod, numpy's transcendental functions
(within a few ulp of the C library's).

Pinned by tests/test_oracle_expr.py: against the compiled reference compiler (oracle/_ref, values and error texts),
and against Python's own evaluation of the same strings (Python's grammar for -, +, *, /, ** and unary signs is the
same as the one above, so precedence and associativity are checked independently of this file).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

MAX_NEST = 200     # cpp/expr.hpp:51
MAX_LEAVES = 2000  # cpp/expr.hpp:52

_libc = C.CDLL(None, use_errno=True)
_libc.strtod.restype = C.c_double
_libc.strtod.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_char))]


def stod(b: bytes):
    """std::stod on bytes: (value, bytes used); ValueError when nothing is read or the value is out of range."""
    buf = C.create_string_buffer(b)
    end = C.POINTER(C.c_char)()
    C.set_errno(0)
    x = _libc.strtod(buf, C.byref(end))
    err = C.get_errno()
    used = C.addressof(end.contents) - C.addressof(buf)
    if used == 0 or err == 34:  # ERANGE
        raise ValueError
    return x, used


class ExprError(RuntimeError):
    """The reference's std::runtime_error of a string that does not compile."""


def _sign(x):
    return (np.asarray(x) > 0).astype(np.float64) - (np.asarray(x) < 0).astype(np.float64)


# cpp/expr.hpp:131-164
FUNCTIONS = {
    "sin": (1, np.sin), "cos": (1, np.cos), "tan": (1, np.tan), "asin": (1, np.arcsin), "acos": (1, np.arccos),
    "atan": (1, np.arctan), "sinh": (1, np.sinh), "cosh": (1, np.cosh), "tanh": (1, np.tanh), "exp": (1, np.exp),
    "log": (1, np.log), "ln": (1, np.log), "log10": (1, np.log10), "sqrt": (1, np.sqrt), "abs": (1, np.fabs),
    "floor": (1, np.floor), "ceil": (1, np.ceil), "sign": (1, _sign),
    "atan2": (2, np.arctan2), "pow": (2, np.power),
    "min": (2, lambda a, b: np.where(a < b, a, b)), "max": (2, lambda a, b: np.where(a > b, a, b)),
    "mod": (2, np.fmod),
}

_SPACE = b" \t\n\v\f\r"  # isspace in the C locale


def _isalpha(c):
    return 65 <= c <= 90 or 97 <= c <= 122


def _isdigit(c):
    return 48 <= c <= 57


class Expr:
    """Compiles text over vars (names bound to the arguments of eval) and params (named constants)."""

    def __init__(self, text, vars, params=None):
        self.src = text.encode("utf-8") if isinstance(text, str) else bytes(text)
        self.vars = list(vars)
        self.params = dict(params or {})
        self.pos = 0
        self.depth = 0
        self.leaves = 0
        self.root = self._sum()
        self._skip()
        if self.pos != len(self.src):
            self._fail("unexpected '" + chr(self.src[self.pos]) + "'")

    def eval(self, *values):
        """The expression at the (broadcast) arrays values[i] of vars[i], as float64."""
        with np.errstate(all="ignore"):
            return np.asarray(self.root([np.asarray(v, dtype=np.float64) for v in values]), dtype=np.float64)

    # ----------------------------------------------------------------------------------------- parser
    def _fail(self, what):
        s = self.src.decode("latin-1")
        shown = s if len(self.src) <= 120 else self.src[:117].decode("latin-1") + "..."
        raise ExprError(f'expression "{shown}": {what} at column {self.pos + 1}')

    def _nest(self):
        self.depth += 1
        if self.depth > MAX_NEST:
            self._fail("expression is nested too deeply")

    def _skip(self):
        while self.pos < len(self.src) and self.src[self.pos] in _SPACE:
            self.pos += 1

    def _eat(self, ch):
        self._skip()
        if self.pos < len(self.src) and self.src[self.pos] == ord(ch):
            self.pos += 1
            return True
        return False

    def _sum(self):
        self._nest()
        left = self._product()
        while True:
            if self._eat("+"):
                a, b = left, self._product()
                left = lambda v, a=a, b=b: a(v) + b(v)
            elif self._eat("-"):
                a, b = left, self._product()
                left = lambda v, a=a, b=b: a(v) - b(v)
            else:
                self.depth -= 1
                return left

    def _product(self):
        left = self._signed()
        while True:
            if self._eat("*"):
                a, b = left, self._signed()
                left = lambda v, a=a, b=b: a(v) * b(v)
            elif self._eat("/"):
                a, b = left, self._signed()
                left = lambda v, a=a, b=b: a(v) / b(v)
            else:
                return left

    def _signed(self):
        self._nest()
        if self._eat("-"):
            a = self._signed()
            out = lambda v, a=a: -a(v)
        elif self._eat("+"):
            out = self._signed()
        else:
            out = self._power()
        self.depth -= 1
        return out

    def _power(self):
        base = self._atom()
        star = False
        if not self._eat("^"):
            self._skip()
            if self.pos + 1 < len(self.src) and self.src[self.pos:self.pos + 2] == b"**":
                self.pos += 2
                star = True
        else:
            star = True
        if star:
            e = self._signed()
            return lambda v, a=base, b=e: np.power(a(v), b(v))
        return base

    def _atom(self):
        self._skip()
        if self.pos >= len(self.src):
            self._fail("unexpected end")
        c = self.src[self.pos]
        self.leaves += 1
        if self.leaves > MAX_LEAVES:
            self._fail("expression has too many terms")
        if c == ord("("):
            self.pos += 1
            e = self._sum()
            if not self._eat(")"):
                self._fail("expected ')'")
            return e
        if _isdigit(c) or c == ord("."):
            try:
                x, used = stod(self.src[self.pos:])
            except ValueError:
                self._fail("number is malformed or out of range")
            self.pos += used
            return lambda v, x=x: np.float64(x)
        if _isalpha(c) or c == ord("_"):
            s = self.pos
            while self.pos < len(self.src) and (_isalpha(self.src[self.pos]) or _isdigit(self.src[self.pos]) or self.src[self.pos] == ord("_")):
                self.pos += 1
            name = self.src[s:self.pos].decode("latin-1")
            self._skip()
            if self.pos < len(self.src) and self.src[self.pos] == ord("("):
                self.pos += 1
                args = []
                if not self._eat(")"):
                    while True:
                        args.append(self._sum())
                        if not self._eat(","):
                            break
                    if not self._eat(")"):
                        self._fail("expected ')'")
                return self._call(name, args)
            if name in self.vars:
                i = self.vars.index(name)
                return lambda v, i=i: v[i]
            if name in self.params:
                x = float(self.params[name])
                return lambda v, x=x: np.float64(x)
            if name == "pi":
                return lambda v: np.float64(3.14159265358979323846)
            if name == "e":
                return lambda v: np.float64(2.71828182845904523536)
            self._fail(f"unknown name '{name}'")
        self._fail("unexpected '" + chr(c) + "'")

    def _call(self, name, args):
        if name not in FUNCTIONS:
            self._fail(f"unknown function '{name}'")
        n, fn = FUNCTIONS[name]
        if len(args) != n:
            self._fail(f"{name} takes {n} argument" + ("" if n == 1 else "s"))
        if n == 1:
            return lambda v, f=fn, a=args[0]: f(a(v))
        return lambda v, f=fn, a=args[0], b=args[1]: f(a(v), b(v))


# ------------------------------------------------------------------------------ the compiled reference compiler
REF_SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "liblmw_ref.so")
_ref_so = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_eval(text, vars, params, x0, x1):
    """The reference's own compiler (cpp/expr.hpp, in oracle/_ref) at the points (x0[i], x1[i]); ExprError with its
    message when the string does not compile."""
    global _ref_so
    if _ref_so is None:
        _ref_so = C.CDLL(REF_SO)
        _ref_so.ref_expr_eval.restype = C.c_int
        _ref_so.ref_last_error.restype = C.c_char_p
    x0 = np.ascontiguousarray(x0, dtype=np.float64).ravel()
    x1 = np.ascontiguousarray(x1, dtype=np.float64).ravel()
    out = np.empty(x0.size)
    names = [k.encode() for k in params]
    pn = (C.c_char_p * max(1, len(names)))(*names)
    pv = (C.c_double * max(1, len(names)))(*params.values())
    dp = C.POINTER(C.c_double)
    rc = _ref_so.ref_expr_eval(text.encode() if isinstance(text, str) else text, vars[0].encode(), vars[1].encode(), len(names), pn, pv,
                               x0.ctypes.data_as(dp), x1.ctypes.data_as(dp), C.c_size_t(x0.size), out.ctypes.data_as(dp))
    if rc != 0:
        raise ExprError(_ref_so.ref_last_error().decode("latin-1"))
    return out


# ------------------------------------------------------------------------------------------ the operators' grids
def sphere_coords(mu, nlon):
    """lon (sphere.cpp:121) and lat (sphere.cpp:120) as the reference forms them (the C library's asin)."""
    lon = np.array([2 * math.pi * k / nlon for k in range(nlon)])
    lat = np.array([math.asin(m) for m in mu])
    return lon, lat


def sphere_params(radius, omega, gravity):
    return {"a": radius, "omega": omega, "g": gravity}  # cpp/main.cpp:162


def sphere_grid(text, mu, nlon, params):
    """One expression over the Gauss grid, (nlat, nlon) south to north."""
    lon, lat = sphere_coords(mu, nlon)
    e = Expr(text, ["lon", "lat"], params)
    return np.broadcast_to(e.eval(lon[None, :], lat[:, None]), (len(mu), nlon)).copy()


def sphere_fields(h, u, v, mu, nlon, params):
    """Sphere.fields (cpp/main.cpp:158-172): the three grids, or ExprError / the non-finite error text."""
    exprs = [Expr(t, ["lon", "lat"], params) for t in (h, u, v)]  # compiled in order, before any evaluation
    lon, lat = sphere_coords(mu, nlon)
    g = [np.broadcast_to(e.eval(lon[None, :], lat[:, None]), (len(mu), nlon)).copy() for e in exprs]
    with np.errstate(all="ignore"):
        if not np.all(np.isfinite(g[0] + g[1] + g[2])):
            raise RuntimeError("Sphere.fields: an expression is not finite on the grid")
    return g


def plane_params(lx, ly, depth, gravity, f0, beta):
    return {"lx": lx, "ly": ly, "H": depth, "g": gravity, "f0": f0, "beta": beta}  # cpp/main.cpp:414


def plane_fields(h, u, v, nx, ny, lx, ly, depth, gravity=9.81, f0=0.0, beta=0.0):
    """Plane.fields (cpp/main.cpp:410-431): h (ny, nx) = H + h(x_center, y_center), u (ny, nx) at (i dx, y_center),
    v (ny + 1, nx) at (x_center, j dy) for j = 1 .. ny-1 and 0 on the walls."""
    par = plane_params(lx, ly, depth, gravity, f0, beta)
    eh, eu, ev = (Expr(t, ["x", "y"], par) for t in (h, u, v))
    dx, dy = lx / nx, ly / ny
    xc = np.array([(i + 0.5) * dx for i in range(nx)])
    xu = np.array([i * dx for i in range(nx)])
    yc = np.array([(j + 0.5) * dy for j in range(ny)])
    yv = np.array([j * dy for j in range(ny + 1)])
    hh = depth + np.broadcast_to(eh.eval(xc[None, :], yc[:, None]), (ny, nx))
    uu = np.broadcast_to(eu.eval(xu[None, :], yc[:, None]), (ny, nx)).copy()
    vv = np.zeros((ny + 1, nx))
    vv[1:ny] = np.broadcast_to(ev.eval(xc[None, :], yv[1:ny, None]), (ny - 1, nx))
    for g, what in ((hh, "height"), (uu, "u"), (vv, "v")):
        if not np.all(np.isfinite(g)):
            raise RuntimeError(f"Plane.fields: the {what} expression is not finite")
    return np.ascontiguousarray(hh), uu, vv
