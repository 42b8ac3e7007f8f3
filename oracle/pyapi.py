"""ctypes bindings of the CPU oracle (liblmw_oracle.so) and of the compiled reference shim (_ref/liblmw_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product package.  Both libraries export the same C API (prefixes ``orc_`` and ``ref_``), so ``OracleSphere`` and
``OraclePlane`` drive either one; they mirror the reference classes lmweather::Sphere (cpp/sphere.hpp:27-121) and
lmweather::Plane (cpp/plane.hpp:20-68) and raise ValueError / RuntimeError where those throw std::invalid_argument /
std::runtime_error, with the same text.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liblmw_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "liblmw_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(ref: bool = True) -> None:
    """Compile the oracle (and, if the reference tree is present, oracle/_ref) with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", HERE, "all" if ref else "liblmw_oracle.so"], check=True)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


_SIGS = {
    "last_error": (C.c_char_p, []),
    "fft": (C.c_int, [_dp, C.c_size_t, C.c_int]),
    "gauss_legendre": (None, [C.c_int, _dp, _dp]),
    "legendre_column": (None, [C.c_int, C.c_int, C.c_double, _dp]),
    "sphere_create": (C.c_void_p, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
    "sphere_destroy": (None, [C.c_void_p]),
    "sphere_dims": (None, [C.c_void_p, _ip, _ip]),
    "sphere_time": (C.c_double, [C.c_void_p]),
    "sphere_gauss": (None, [C.c_void_p, _dp, _dp]),
    "sphere_set_diffusion": (None, [C.c_void_p, C.c_double]),
    "sphere_set_advect_only": (None, [C.c_void_p, C.c_int]),
    "sphere_set_axis_tilt": (None, [C.c_void_p, C.c_double]),
    "sphere_set_orography_grid": (C.c_int, [C.c_void_p, _dp, C.c_size_t]),
    "sphere_mark_reference": (None, [C.c_void_p]),
    "sphere_set_state_grid": (C.c_int, [C.c_void_p, _dp, _dp, _dp, C.c_size_t]),
    "sphere_set_tracer_grid": (C.c_int, [C.c_void_p, _dp, C.c_size_t]),
    "sphere_add_mode": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double]),
    "sphere_perturb": (C.c_int, [C.c_void_p, C.c_double, C.c_uint64, C.c_int]),
    "sphere_step": (None, [C.c_void_p, C.c_double]),
    "sphere_run": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "sphere_stable_dt": (C.c_int, [C.c_void_p, C.c_double, _dp]),
    "sphere_grid_height": (None, [C.c_void_p, _dp]),
    "sphere_grid_vorticity": (None, [C.c_void_p, _dp]),
    "sphere_grid_u": (None, [C.c_void_p, _dp]),
    "sphere_grid_v": (None, [C.c_void_p, _dp]),
    "sphere_grid_tracer": (C.c_int, [C.c_void_p, _dp]),
    "sphere_mean_height": (C.c_double, [C.c_void_p]),
    "sphere_kinetic_energy": (C.c_double, [C.c_void_p]),
    "sphere_potential_energy": (C.c_double, [C.c_void_p]),
    "sphere_energy": (C.c_double, [C.c_void_p]),
    "sphere_spectrum": (None, [C.c_void_p, _dp, _dp]),
    "sphere_potential_enstrophy": (C.c_double, [C.c_void_p]),
    "sphere_max_wind": (C.c_double, [C.c_void_p]),
    "sphere_height_errors": (None, [C.c_void_p, _dp, _dp]),
    "sphere_tracer_errors": (C.c_int, [C.c_void_p, _dp, _dp]),
    "sphere_tracer_mean": (C.c_int, [C.c_void_p, _dp]),
    "sphere_tracer_mass": (C.c_int, [C.c_void_p, _dp]),
    "sphere_height_at": (C.c_double, [C.c_void_p, C.c_double, C.c_double]),
    "sphere_vorticity_at": (C.c_double, [C.c_void_p, C.c_double, C.c_double]),
    "sphere_tracer_at": (C.c_int, [C.c_void_p, C.c_double, C.c_double, _dp]),
    "sphere_wind_at": (C.c_double, [C.c_void_p, C.c_double, C.c_double, C.c_int]),
    "sphere_mode_coefficient": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp]),
    "sphere_wave_crest": (C.c_int, [C.c_void_p, C.c_int, C.c_double, _dp]),
    "sphere_synthesize": (None, [C.c_void_p, _dp, _dp]),
    "sphere_analyze": (None, [C.c_void_p, _dp, _dp]),
    "sphere_analyze_divergence": (None, [C.c_void_p, _dp, _dp, _dp]),
    "sphere_winds": (None, [C.c_void_p, _dp, _dp, _dp, _dp]),
    "sphere_tendencies": (None, [C.c_void_p, _dp, _dp, _dp]),
    "sphere_get_spectral": (None, [C.c_void_p, C.c_int, _dp]),
    "sphere_set_spectral": (None, [C.c_void_p, C.c_int, _dp]),
    "sphere_get_tables": (None, [C.c_void_p, _dp, _dp]),
    "sphere_get_coriolis": (None, [C.c_void_p, _dp]),
    "plane_create": (C.c_void_p, [C.c_int, C.c_int] + [C.c_double] * 8 + [C.c_int]),
    "plane_destroy": (None, [C.c_void_p]),
    "plane_set_params": (None, [C.c_void_p] + [C.c_double] * 6 + [C.c_int]),
    "plane_time": (C.c_double, [C.c_void_p]),
    "plane_reset": (None, [C.c_void_p]),
    "plane_perturb": (None, [C.c_void_p, C.c_double, C.c_uint64]),
    "plane_step": (None, [C.c_void_p, C.c_double]),
    "plane_run": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "plane_stable_dt": (C.c_int, [C.c_void_p, C.c_double, _dp]),
    "plane_mean_depth": (C.c_double, [C.c_void_p]),
    "plane_energy": (C.c_double, [C.c_void_p]),
    "plane_max_speed": (C.c_double, [C.c_void_p]),
    "plane_pv_anomaly_rms": (C.c_double, [C.c_void_p]),
    "plane_zeta_at_corner": (C.c_double, [C.c_void_p, C.c_int, C.c_int]),
    "plane_at": (C.c_double, [C.c_void_p, C.c_double, C.c_double, C.c_int]),
    "plane_zonal_amplitude": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
    "plane_zonal_crest": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
    "plane_harmonic_energy": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "plane_v_zonal_crest": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
    "plane_set_field": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_size_t]),
    "plane_get_field": (None, [C.c_void_p, C.c_int, _dp]),
    "plane_tendency": (None, [C.c_void_p, _dp, _dp, _dp]),
}

class Lib:
    """One of the two CPU libraries (``kind`` = "oracle" or "ref")."""

    def __init__(self, kind: str = "oracle"):
        self.kind = kind
        path = ORACLE_SO if kind == "oracle" else REF_SO
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.so = C.CDLL(path)
        prefix = "orc_" if kind == "oracle" else "ref_"
        for name, (res, args) in _SIGS.items():
            fn = getattr(self.so, prefix + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def check(self, rc: int) -> None:
        if rc == 0:
            return
        msg = self.last_error().decode()
        raise (ValueError if rc == 1 else RuntimeError)(msg)


def available(kind: str) -> bool:
    return os.path.exists(ORACLE_SO if kind == "oracle" else REF_SO)


class OracleSphere:
    """lmweather::Sphere (cpp/sphere.hpp:27-121) on one CPU core, through either library."""

    def __init__(self, truncation: int, radius=6.37122e6, omega=7.292e-5, gravity=9.80616, diffusion=0.0, kind="oracle"):
        self.lib = Lib(kind) if isinstance(kind, str) else kind
        self.h = self.lib.sphere_create(truncation, radius, omega, gravity, diffusion)
        if not self.h:
            raise ValueError(self.lib.last_error().decode())
        self.N = truncation
        self.radius, self.omega, self.gravity = radius, omega, gravity
        a, b = C.c_int(), C.c_int()
        self.lib.sphere_dims(self.h, C.byref(a), C.byref(b))
        self.nlat, self.nlon = a.value, b.value
        self.ngrid = self.nlat * self.nlon
        self.nspec = (truncation + 1) ** 2

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.sphere_destroy(self.h)
            self.h = None

    # set-up
    def set_state_grid(self, h, u, v):
        h, u, v = (np.ascontiguousarray(x, dtype=np.float64).ravel() for x in (h, u, v))
        self.lib.check(self.lib.sphere_set_state_grid(self.h, _ptr(h), _ptr(u), _ptr(v), h.size if h.size == u.size == v.size else 0))

    def set_orography_grid(self, hs):
        hs = np.ascontiguousarray(hs, dtype=np.float64).ravel()
        self.lib.check(self.lib.sphere_set_orography_grid(self.h, _ptr(hs), hs.size))

    def set_tracer_grid(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64).ravel()
        self.lib.check(self.lib.sphere_set_tracer_grid(self.h, _ptr(q), q.size))

    def set_axis_tilt(self, alpha):
        self.lib.sphere_set_axis_tilt(self.h, alpha)

    def set_diffusion(self, nu4):
        self.lib.sphere_set_diffusion(self.h, nu4)

    def set_advect_only(self, on):
        self.lib.sphere_set_advect_only(self.h, int(bool(on)))

    def mark_reference(self):
        self.lib.sphere_mark_reference(self.h)

    def add_mode(self, n, m, amplitude, mean_depth):
        self.lib.check(self.lib.sphere_add_mode(self.h, n, m, amplitude, mean_depth))

    def perturb(self, amplitude, seed, max_wave):
        self.lib.check(self.lib.sphere_perturb(self.h, amplitude, seed, max_wave))

    # stepping
    def step(self, dt):
        self.lib.sphere_step(self.h, dt)

    def run(self, seconds, dt):
        self.lib.check(self.lib.sphere_run(self.h, seconds, dt))

    def stable_dt(self, safety=1.5):
        out = C.c_double()
        self.lib.check(self.lib.sphere_stable_dt(self.h, safety, C.byref(out)))
        return out.value

    def time(self):
        return self.lib.sphere_time(self.h)

    # grids
    def _grid(self, fn):
        g = np.empty(self.ngrid)
        r = fn(self.h, _ptr(g))
        if r:
            self.lib.check(r)
        return g.reshape(self.nlat, self.nlon)

    def grid_height(self):
        return self._grid(self.lib.sphere_grid_height)

    def grid_vorticity(self):
        return self._grid(self.lib.sphere_grid_vorticity)

    def grid_u(self):
        return self._grid(self.lib.sphere_grid_u)

    def grid_v(self):
        return self._grid(self.lib.sphere_grid_v)

    def grid_tracer(self):
        return self._grid(self.lib.sphere_grid_tracer)

    def gauss(self):
        mu, w = np.empty(self.nlat), np.empty(self.nlat)
        self.lib.sphere_gauss(self.h, _ptr(mu), _ptr(w))
        return mu, w

    def coriolis(self):
        return self._grid(self.lib.sphere_get_coriolis)

    # diagnostics
    def mean_height(self):
        return self.lib.sphere_mean_height(self.h)

    def energy(self):
        return self.lib.sphere_energy(self.h)

    def kinetic_energy(self):
        return self.lib.sphere_kinetic_energy(self.h)

    def potential_energy(self):
        return self.lib.sphere_potential_energy(self.h)

    def potential_enstrophy(self):
        return self.lib.sphere_potential_enstrophy(self.h)

    def max_wind(self):
        return self.lib.sphere_max_wind(self.h)

    def spectrum(self):
        ke, pe = np.empty(self.N + 1), np.empty(self.N + 1)
        self.lib.sphere_spectrum(self.h, _ptr(ke), _ptr(pe))
        return ke, pe

    def height_errors(self):
        a, b = C.c_double(), C.c_double()
        self.lib.sphere_height_errors(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def tracer_errors(self):
        a, b = C.c_double(), C.c_double()
        self.lib.check(self.lib.sphere_tracer_errors(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def _scalar(self, fn, *args):
        out = C.c_double()
        self.lib.check(fn(self.h, *args, C.byref(out)))
        return out.value

    def tracer_mean(self):
        return self._scalar(self.lib.sphere_tracer_mean)

    def tracer_mass(self):
        return self._scalar(self.lib.sphere_tracer_mass)

    def tracer_at(self, lat, lon):
        return self._scalar(self.lib.sphere_tracer_at, lat, lon)

    def height_at(self, lat, lon):
        return self.lib.sphere_height_at(self.h, lat, lon)

    def vorticity_at(self, lat, lon):
        return self.lib.sphere_vorticity_at(self.h, lat, lon)

    def wind_at(self, lat, lon, zonal):
        return self.lib.sphere_wind_at(self.h, lat, lon, int(bool(zonal)))

    def mode_coefficient(self, n, m, imag):
        return self._scalar(self.lib.sphere_mode_coefficient, n, m, int(bool(imag)))

    def wave_crest(self, m, lat):
        return self._scalar(self.lib.sphere_wave_crest, m, lat)

    # transforms and raw state (complex arrays of shape (N+1, N+1), index [m, n])
    def synthesize(self, spec):
        s = np.ascontiguousarray(spec, dtype=np.complex128).ravel()
        g = np.empty(self.ngrid)
        self.lib.sphere_synthesize(self.h, s.ctypes.data_as(_dp), _ptr(g))
        return g.reshape(self.nlat, self.nlon)

    def analyze(self, grid):
        g = np.ascontiguousarray(grid, dtype=np.float64).ravel()
        s = np.empty(self.nspec, dtype=np.complex128)
        self.lib.sphere_analyze(self.h, _ptr(g), s.ctypes.data_as(_dp))
        return s.reshape(self.N + 1, self.N + 1)

    def analyze_divergence(self, A, B):
        A = np.ascontiguousarray(A, dtype=np.float64).ravel()
        B = np.ascontiguousarray(B, dtype=np.float64).ravel()
        s = np.empty(self.nspec, dtype=np.complex128)
        self.lib.sphere_analyze_divergence(self.h, _ptr(A), _ptr(B), s.ctypes.data_as(_dp))
        return s.reshape(self.N + 1, self.N + 1)

    def winds(self, z, d):
        z = np.ascontiguousarray(z, dtype=np.complex128).ravel()
        d = np.ascontiguousarray(d, dtype=np.complex128).ravel()
        U, V = np.empty(self.ngrid), np.empty(self.ngrid)
        self.lib.sphere_winds(self.h, z.ctypes.data_as(_dp), d.ctypes.data_as(_dp), _ptr(U), _ptr(V))
        return U.reshape(self.nlat, self.nlon), V.reshape(self.nlat, self.nlon)

    def tendencies(self):
        out = [np.empty(self.nspec, dtype=np.complex128) for _ in range(3)]
        self.lib.sphere_tendencies(self.h, *(o.ctypes.data_as(_dp) for o in out))
        return [o.reshape(self.N + 1, self.N + 1) for o in out]

    def get_spectral(self, field):
        s = np.empty(self.nspec, dtype=np.complex128)
        self.lib.sphere_get_spectral(self.h, field, s.ctypes.data_as(_dp))
        return s.reshape(self.N + 1, self.N + 1)

    def set_spectral(self, field, spec):
        s = np.ascontiguousarray(spec, dtype=np.complex128).ravel()
        self.lib.sphere_set_spectral(self.h, field, s.ctypes.data_as(_dp))

    def tables(self):
        n = self.nlat * (self.N + 1) ** 2
        P, H = np.empty(n), np.empty(n)
        self.lib.sphere_get_tables(self.h, _ptr(P), _ptr(H))
        shp = (self.nlat, self.N + 1, self.N + 1)
        return P.reshape(shp), H.reshape(shp)


class OraclePlane:
    """lmweather::Plane (cpp/plane.hpp:20-68) on one CPU core, through either library."""

    def __init__(self, nx, ny, lx, ly, gravity=9.81, depth=100.0, f0=0.0, beta=0.0, viscosity=0.0, drag=0.0, linear=False,
                 kind="oracle"):
        self.lib = Lib(kind) if isinstance(kind, str) else kind
        self.h = self.lib.plane_create(nx, ny, lx, ly, gravity, depth, f0, beta, viscosity, drag, int(bool(linear)))
        if not self.h:
            raise ValueError(self.lib.last_error().decode())
        self.nx, self.ny, self.lx, self.ly = nx, ny, lx, ly

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.plane_destroy(self.h)
            self.h = None

    def set_params(self, gravity, depth, f0, beta, viscosity, drag, linear):
        self.lib.plane_set_params(self.h, gravity, depth, f0, beta, viscosity, drag, int(bool(linear)))

    def reset(self):
        self.lib.plane_reset(self.h)

    def perturb(self, amplitude, seed):
        self.lib.plane_perturb(self.h, amplitude, seed)

    def step(self, dt):
        self.lib.plane_step(self.h, dt)

    def run(self, seconds, dt):
        self.lib.check(self.lib.plane_run(self.h, seconds, dt))

    def stable_dt(self, safety=0.5):
        out = C.c_double()
        self.lib.check(self.lib.plane_stable_dt(self.h, safety, C.byref(out)))
        return out.value

    def time(self):
        return self.lib.plane_time(self.h)

    def set_field(self, field, g):
        g = np.ascontiguousarray(g, dtype=np.float64).ravel()
        self.lib.check(self.lib.plane_set_field(self.h, field, _ptr(g), g.size))

    def get_field(self, field):
        rows = self.ny + 1 if field == 2 else self.ny
        g = np.empty(rows * self.nx)
        self.lib.plane_get_field(self.h, field, _ptr(g))
        return g.reshape(rows, self.nx)

    def tendency(self):
        dh, du, dv = np.empty(self.nx * self.ny), np.empty(self.nx * self.ny), np.empty(self.nx * (self.ny + 1))
        self.lib.plane_tendency(self.h, _ptr(dh), _ptr(du), _ptr(dv))
        return dh.reshape(self.ny, self.nx), du.reshape(self.ny, self.nx), dv.reshape(self.ny + 1, self.nx)

    def mean_depth(self):
        return self.lib.plane_mean_depth(self.h)

    def energy(self):
        return self.lib.plane_energy(self.h)

    def max_speed(self):
        return self.lib.plane_max_speed(self.h)

    def pv_anomaly_rms(self):
        return self.lib.plane_pv_anomaly_rms(self.h)

    def zeta_at_corner(self, i, j):
        return self.lib.plane_zeta_at_corner(self.h, i, j)

    def at(self, x, y, kind):
        return self.lib.plane_at(self.h, x, y, kind)

    def _scalar(self, fn, *args):
        out = C.c_double()
        self.lib.check(fn(self.h, *args, C.byref(out)))
        return out.value

    def zonal_amplitude(self, m, j):
        return self._scalar(self.lib.plane_zonal_amplitude, m, j)

    def zonal_crest(self, m, j):
        return self._scalar(self.lib.plane_zonal_crest, m, j)

    def harmonic_energy(self, m):
        return self._scalar(self.lib.plane_harmonic_energy, m)

    def v_zonal_crest(self, m, j):
        return self._scalar(self.lib.plane_v_zonal_crest, m, j)


def fft(re_im: np.ndarray, sign: int, kind="oracle") -> np.ndarray:
    lib = Lib(kind)
    a = np.ascontiguousarray(re_im, dtype=np.complex128).copy()
    lib.check(lib.fft(a.ctypes.data_as(_dp), a.size, sign))
    return a


def gauss_legendre(n: int, kind="oracle"):
    lib = Lib(kind)
    x, w = np.empty(n), np.empty(n)
    lib.gauss_legendre(n, _ptr(x), _ptr(w))
    return x, w


def legendre_column(m: int, nmax: int, mu: float, kind="oracle") -> np.ndarray:
    lib = Lib(kind)
    out = np.empty(nmax + 2 - m)
    lib.legendre_column(m, nmax, mu, _ptr(out))
    return out
