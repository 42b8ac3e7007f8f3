"""Host mirror of lmweather::Sphere (reference cpp/sphere.hpp:27-121) over the gfx950 C ABI.

Same method names, argument meanings and error behaviour as the reference class: ValueError where it throws
std::invalid_argument, RuntimeError where it throws std::runtime_error, with identical messages.  State lives on the
GPU; grids are (nlat, nlon) float64 arrays (south to north) and spectral arrays are complex (N+1, N+1) arrays indexed
[m, n] like the reference's m*(N+1)+n storage.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import _lib as L


class Sphere:
    """One spectral shallow-water model (reference Sphere) resident on the GPU."""

    def __init__(self, truncation: int, params: L.SphereParams | None = None, precision: int = L.FP64):
        self.params = params if params is not None else L.SphereParams()
        h = C.c_void_p()
        L.call("lmw_sphere_create", truncation, C.byref(self.params), C.byref(h))
        self._h = h
        t, nlat, nlon = C.c_int(), C.c_int(), C.c_int()
        L.call("lmw_sphere_dims", self._h, C.byref(t), C.byref(nlat), C.byref(nlon))
        self.N, self._nlat, self._nlon = t.value, nlat.value, nlon.value
        self._mu = self._w = None
        if precision != L.FP64:
            self.set_precision(precision)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                L.lib().lmw_sphere_destroy(h)
            except Exception:
                pass
            self._h = None

    # ---------------------------------------------------------------- geometry (sphere.hpp:30-36)
    def truncation(self) -> int:
        return self.N

    def nlat(self) -> int:
        return self._nlat

    def nlon(self) -> int:
        return self._nlon

    def time(self) -> float:
        return self._diag(L.DIAG_TIME)

    def _gauss(self):
        if self._mu is None:
            mu, w = np.empty(self._nlat), np.empty(self._nlat)
            L.call("lmw_sphere_gauss", self._h, L.dptr(mu), L.dptr(w))
            self._mu, self._w = mu, w
        return self._mu, self._w

    def gauss_mu(self) -> np.ndarray:
        return self._gauss()[0].copy()

    def gauss_weights(self) -> np.ndarray:
        return self._gauss()[1].copy()

    def lat(self, j: int) -> float:  # sphere.cpp:120
        return math.asin(self._gauss()[0][j])

    def lon(self, k: int) -> float:  # sphere.cpp:121
        return 2 * math.pi * k / self._nlon

    def lats(self) -> np.ndarray:
        return np.arcsin(self._gauss()[0])

    def lons(self) -> np.ndarray:
        return 2 * np.pi * np.arange(self._nlon) / self._nlon

    @property
    def ngrid(self) -> int:
        return self._nlat * self._nlon

    # ---------------------------------------------------------------- set-up
    def set_precision(self, precision: int) -> None:
        L.call("lmw_sphere_set_precision", self._h, precision)

    def set_state_grid(self, h, u, v) -> None:
        h, u, v = (L.as_f64(x).ravel() for x in (h, u, v))
        n = h.size if h.size == u.size == v.size else 0
        L.call("lmw_sphere_set_state_grid", self._h, L.dptr(h), L.dptr(u), L.dptr(v), n)

    def set_orography_grid(self, hs) -> None:
        hs = L.as_f64(hs).ravel()
        L.call("lmw_sphere_set_orography_grid", self._h, L.dptr(hs), hs.size)

    def set_axis_tilt(self, alpha: float) -> None:
        L.call("lmw_sphere_set_axis_tilt", self._h, alpha)

    def set_diffusion(self, nu4: float) -> None:
        L.call("lmw_sphere_set_diffusion", self._h, nu4)

    def mark_reference(self) -> None:
        L.call("lmw_sphere_mark_reference", self._h)

    # ---------------------------------------------------------------- tracer (sphere.hpp:52-64)
    def set_tracer_grid(self, q) -> None:
        g = L.as_f64(q).ravel()
        L.call("lmw_sphere_set_tracer_grid", self._h, L.dptr(g), g.size)

    def set_advect_only(self, on: bool) -> None:
        L.call("lmw_sphere_set_advect_only", self._h, int(bool(on)))

    def _flags(self):
        a, b = C.c_int(), C.c_int()
        L.call("lmw_sphere_tracer_flags", self._h, C.byref(a), C.byref(b))
        return bool(a.value), bool(b.value)

    def has_tracer(self) -> bool:
        return self._flags()[0]

    def advect_only(self) -> bool:
        return self._flags()[1]

    def grid_tracer(self) -> np.ndarray:
        return self._grid(L.GRID_TRACER)

    def tracer_mean(self) -> float:
        return self._diag(L.DIAG_TRACER_MEAN)

    def tracer_mass(self) -> float:
        return self._diag(L.DIAG_TRACER_MASS)

    def tracer_min(self) -> float:
        return self._diag(L.DIAG_TRACER_MIN)

    def tracer_max(self) -> float:
        return self._diag(L.DIAG_TRACER_MAX)

    def tracer_error_l2(self) -> float:
        return self._diag(L.DIAG_TRACER_ERROR_L2)

    def tracer_error_linf(self) -> float:
        return self._diag(L.DIAG_TRACER_ERROR_LINF)

    def tracer_at(self, lat: float, lon: float) -> float:
        return self._point(L.POINT_TRACER, lat, lon)

    def autotune(self, force: bool = False):
        """Choose the Legendre tile shapes for this sphere (timed on its own state, which is restored, or taken from
        the tile cache); returns (tiles[4], best_ms[2])."""
        tiles = np.zeros(4, dtype=np.int32)
        ms = np.zeros(2)
        L.call("lmw_sphere_autotune", self._h, int(bool(force)), tiles.ctypes.data_as(L._ip), L.dptr(ms))
        return tiles, ms

    def setup(self, which: int, alpha: float = 0.0) -> None:
        """A Williamson et al. (1992) test case (L.SETUP_WILLIAMSON1/2/5/6: the williamson1(alpha), williamson2(alpha),
        williamson5(), williamson6() operators of main.cpp:174-291), evaluated and analysed on the device."""
        L.call("lmw_sphere_setup", self._h, int(which), float(alpha))

    def fields(self, h: str, u: str, v: str) -> None:
        """The Sphere.fields operator (main.cpp:158-172): h, u, v from expressions in lon, lat and a, omega, g
        (cpp/expr.hpp), evaluated at the Gauss grid on the device, then set_state_grid."""
        L.call("lmw_sphere_fields", self._h, L.text(h), L.text(u), L.text(v))

    def orography(self, hs: str) -> None:
        """The Sphere.orography operator (main.cpp:212-228): set_orography_grid of an expression in lon, lat."""
        L.call("lmw_sphere_orography", self._h, L.text(hs))

    def tracer(self, q: str) -> None:
        """The Sphere.tracer operator (main.cpp:252-268): set_tracer_grid of an expression in lon, lat."""
        L.call("lmw_sphere_tracer", self._h, L.text(q))

    def add_mode(self, n: int, m: int, amplitude: float, mean_depth: float) -> None:
        L.call("lmw_sphere_add_mode", self._h, n, m, amplitude, mean_depth)

    def perturb(self, amplitude: float, seed: int, max_wave: int) -> None:
        L.call("lmw_sphere_perturb", self._h, amplitude, seed, max_wave)

    # ---------------------------------------------------------------- stepping
    def step(self, dt: float) -> None:
        L.call("lmw_sphere_step", self._h, dt)

    def steps(self, dt: float, count: int) -> None:
        """`count` steps of dt with no periodic checks (replays of a HIP graph of 8 steps, then of a one-step graph)."""
        L.call("lmw_sphere_steps", self._h, dt, count)

    def run(self, seconds: float, dt: float) -> None:
        L.call("lmw_sphere_run", self._h, seconds, dt)

    def stable_dt(self, safety: float = 1.5) -> float:
        out = C.c_double()
        L.call("lmw_sphere_stable_dt", self._h, safety, C.byref(out))
        return out.value

    # ---------------------------------------------------------------- diagnostics
    def _diag(self, which: int) -> float:
        out = C.c_double()
        L.call("lmw_sphere_diag", self._h, which, C.byref(out))
        return out.value

    def mean_height(self) -> float:
        return self._diag(L.DIAG_MEAN_HEIGHT)

    def energy(self) -> float:
        return self._diag(L.DIAG_ENERGY)

    def kinetic_energy(self) -> float:
        return self._diag(L.DIAG_KINETIC)

    def potential_energy(self) -> float:
        return self._diag(L.DIAG_POTENTIAL)

    def potential_enstrophy(self) -> float:
        return self._diag(L.DIAG_ENSTROPHY)

    def max_wind(self) -> float:
        return self._diag(L.DIAG_MAX_WIND)

    def height_error_l2(self) -> float:
        return self._diag(L.DIAG_ERROR_L2)

    def height_error_linf(self) -> float:
        return self._diag(L.DIAG_ERROR_LINF)

    def spectrum(self):
        ke, pe = np.empty(self.N + 1), np.empty(self.N + 1)
        L.call("lmw_sphere_spectrum", self._h, L.dptr(ke), L.dptr(pe), self.N + 1)
        return ke, pe

    def _point(self, which, lat, lon) -> float:
        out = C.c_double()
        L.call("lmw_sphere_point", self._h, which, lat, lon, C.byref(out))
        return out.value

    def height_at(self, lat: float, lon: float) -> float:
        return self._point(L.POINT_HEIGHT, lat, lon)

    def wind_at(self, lat: float, lon: float, zonal: bool) -> float:
        return self._point(L.POINT_U if zonal else L.POINT_V, lat, lon)

    def vorticity_at(self, lat: float, lon: float) -> float:
        return self._point(L.POINT_VORTICITY, lat, lon)

    def mode_coefficient(self, n: int, m: int, imag: bool) -> float:
        out = C.c_double()
        L.call("lmw_sphere_mode_coefficient", self._h, n, m, int(bool(imag)), C.byref(out))
        return out.value

    def wave_crest(self, m: int, lat: float) -> float:
        out = C.c_double()
        L.call("lmw_sphere_wave_crest", self._h, m, lat, C.byref(out))
        return out.value

    def _grid(self, field: int) -> np.ndarray:
        g = np.empty(self.ngrid)
        L.call("lmw_sphere_get_grid", self._h, field, L.dptr(g), g.size)
        return g.reshape(self._nlat, self._nlon)

    def grid_height(self) -> np.ndarray:
        return self._grid(L.GRID_H)

    def grid_u(self) -> np.ndarray:
        return self._grid(L.GRID_U)

    def grid_v(self) -> np.ndarray:
        return self._grid(L.GRID_V)

    def grid_vorticity(self) -> np.ndarray:
        return self._grid(L.GRID_VORTICITY)

    def grid_orography(self) -> np.ndarray:  # sphere.hpp:50
        return self._grid(L.GRID_OROGRAPHY)

    # ---------------------------------------------------------------- text output (main.cpp:328-383)
    def save_csv(self, path: str) -> None:
        """Rows "lon lat h u v zeta q hs" (degrees, "%.10g"), a blank line after each latitude row."""
        L.call("lmw_sphere_save_csv", self._h, os.fsencode(path))

    def save_spectrum(self, path: str) -> None:
        """Rows "n ke pe" for n = 1..T."""
        L.call("lmw_sphere_save_spectrum", self._h, os.fsencode(path))

    def save_zonal_mean(self, path: str) -> None:
        """Rows "lat mean_h mean_u" per latitude."""
        L.call("lmw_sphere_save_zonal_mean", self._h, os.fsencode(path))

    # ---------------------------------------------------------------- spectral state and transforms
    def get_spectral(self, field: int) -> np.ndarray:
        s = np.empty((self.N + 1) ** 2, dtype=np.complex128)
        L.call("lmw_sphere_get_spectral", self._h, field, s.ctypes.data_as(L._dp), s.size)
        return s.reshape(self.N + 1, self.N + 1)

    def set_spectral(self, field: int, spec) -> None:
        s = L.as_c128(spec).ravel()
        L.call("lmw_sphere_set_spectral", self._h, field, s.ctypes.data_as(L._dp), s.size)

    def synthesize_scalar(self, spec) -> np.ndarray:
        s = L.as_c128(spec).ravel()
        g = np.empty(self.ngrid)
        L.call("lmw_sphere_synthesize", self._h, s.ctypes.data_as(L._dp), L.dptr(g))
        return g.reshape(self._nlat, self._nlon)

    def analyze_scalar(self, grid) -> np.ndarray:
        g = L.as_f64(grid).ravel()
        s = np.empty((self.N + 1) ** 2, dtype=np.complex128)
        L.call("lmw_sphere_analyze", self._h, L.dptr(g), s.ctypes.data_as(L._dp))
        return s.reshape(self.N + 1, self.N + 1)

    def analyze_divergence(self, A, B) -> np.ndarray:
        A, B = L.as_f64(A).ravel(), L.as_f64(B).ravel()
        s = np.empty((self.N + 1) ** 2, dtype=np.complex128)
        L.call("lmw_sphere_analyze_divergence", self._h, L.dptr(A), L.dptr(B), s.ctypes.data_as(L._dp))
        return s.reshape(self.N + 1, self.N + 1)

    def winds_spectral_to_grid(self, zeta, delta):
        z, d = L.as_c128(zeta).ravel(), L.as_c128(delta).ravel()
        U, V = np.empty(self.ngrid), np.empty(self.ngrid)
        L.call("lmw_sphere_winds", self._h, z.ctypes.data_as(L._dp), d.ctypes.data_as(L._dp), L.dptr(U), L.dptr(V))
        return U.reshape(self._nlat, self._nlon), V.reshape(self._nlat, self._nlon)

    def tendencies(self):
        out = [np.empty((self.N + 1) ** 2, dtype=np.complex128) for _ in range(3)]
        L.call("lmw_sphere_tendencies", self._h, *(o.ctypes.data_as(L._dp) for o in out))
        return [o.reshape(self.N + 1, self.N + 1) for o in out]

    def tables(self):
        n = self._nlat * (self.N + 1) ** 2
        P, H = np.empty(n), np.empty(n)
        L.call("lmw_sphere_get_tables", self._h, L.dptr(P), L.dptr(H))
        shp = (self._nlat, self.N + 1, self.N + 1)
        return P.reshape(shp), H.reshape(shp)


class Ensemble:
    """Independent Sphere members of possibly different truncations, stepped together on the GPU."""

    def __init__(self, truncations, params: L.SphereParams | None = None, precision: int = L.FP64):
        self.params = params if params is not None else L.SphereParams()
        tr = np.ascontiguousarray(truncations, dtype=np.int32)
        h = C.c_void_p()
        L.call("lmw_ensemble_create", int(tr.size), tr.ctypes.data_as(L._ip), C.byref(self.params), precision, C.byref(h))
        self._h = h
        self.truncations = [int(x) for x in tr]
        ng = C.c_int()
        L.call("lmw_ensemble_groups", self._h, C.byref(ng), None, None)
        gt, gm = (C.c_int * ng.value)(), (C.c_int * ng.value)()
        L.call("lmw_ensemble_groups", self._h, C.byref(ng), gt, gm)
        self.group_truncations = list(gt)
        self.group_members = list(gm)
        self._dims = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                L.lib().lmw_ensemble_destroy(h)
            except Exception:
                pass
            self._h = None

    def __len__(self):
        return len(self.truncations)

    def dims(self, member: int):
        if member not in self._dims:
            t, a, b = C.c_int(), C.c_int(), C.c_int()
            L.call("lmw_ensemble_member_dims", self._h, member, C.byref(t), C.byref(a), C.byref(b))
            self._dims[member] = (t.value, a.value, b.value)
        return self._dims[member]

    def set_state_grid(self, member, h, u, v) -> None:
        h, u, v = (L.as_f64(x).ravel() for x in (h, u, v))
        L.call("lmw_ensemble_set_state_grid", self._h, member, L.dptr(h), L.dptr(u), L.dptr(v), h.size)

    def set_state_grids(self, member0: int, count: int, h, u, v) -> None:
        """set_state_grid of members member0 .. member0 + count - 1 (one truncation) from (count, nlat, nlon) arrays,
        in batched launches."""
        h, u, v = (L.as_f64(x).ravel() for x in (h, u, v))
        L.call("lmw_ensemble_set_state_grids", self._h, int(member0), int(count), L.dptr(h), L.dptr(u), L.dptr(v), h.size)

    def setup(self, member0: int, count: int, which: int, alpha: float = 0.0) -> None:
        """Sphere.setup for members member0 .. member0 + count - 1 (test case 1 needs a tracer: single spheres only)."""
        L.call("lmw_ensemble_setup", self._h, int(member0), int(count), int(which), float(alpha))

    def fields(self, member0: int, count: int, h: str, u: str, v: str) -> None:
        """Sphere.fields for members member0 .. member0 + count - 1 (one grid per truncation group)."""
        L.call("lmw_ensemble_fields", self._h, int(member0), int(count), L.text(h), L.text(u), L.text(v))

    def orography(self, member0: int, count: int, hs: str) -> None:
        """Sphere.orography for members member0 .. member0 + count - 1 (one shared plane per truncation group)."""
        L.call("lmw_ensemble_orography", self._h, int(member0), int(count), L.text(hs))

    def set_orography_grid(self, member, hs) -> None:
        hs = L.as_f64(hs).ravel()
        L.call("lmw_ensemble_set_orography_grid", self._h, member, L.dptr(hs), hs.size)

    def set_axis_tilt(self, member, alpha) -> None:
        L.call("lmw_ensemble_set_axis_tilt", self._h, member, alpha)

    def set_spectral(self, member, field, spec) -> None:
        s = L.as_c128(spec).ravel()
        L.call("lmw_ensemble_set_spectral", self._h, member, field, s.ctypes.data_as(L._dp), s.size)

    def get_spectral(self, member, field) -> np.ndarray:
        N = self.dims(member)[0]
        s = np.empty((N + 1) ** 2, dtype=np.complex128)
        L.call("lmw_ensemble_get_spectral", self._h, member, field, s.ctypes.data_as(L._dp), s.size)
        return s.reshape(N + 1, N + 1)

    def grid(self, member, field) -> np.ndarray:
        _, nlat, nlon = self.dims(member)
        g = np.empty(nlat * nlon)
        L.call("lmw_ensemble_get_grid", self._h, member, field, L.dptr(g), g.size)
        return g.reshape(nlat, nlon)

    def diag(self, member, which) -> float:
        out = C.c_double()
        L.call("lmw_ensemble_diag", self._h, member, which, C.byref(out))
        return out.value

    def get_truncated(self, truncation: int) -> np.ndarray:
        """Spectral states of all members (single-truncation ensemble) truncated to ``truncation``: float32 array
        (members, 3, C, 2) in the packed triangular order, C = (T+1)(T+2)/2."""
        T = int(truncation)
        out = np.empty((len(self.truncations), 3, (T + 1) * (T + 2) // 2, 2), dtype=np.float32)
        L.call("lmw_ensemble_get_truncated", self._h, T, L.fptr(out), out.size)
        return out

    def set_truncated(self, truncation: int, states) -> None:
        """Set every member's spectral state from a (members, 3, C_T, 2) float32 array (zero above T)."""
        a = L.as_f32(states)
        L.call("lmw_ensemble_set_truncated", self._h, int(truncation), L.fptr(a), a.size)

    def steps(self, dts, counts, graph: bool = True) -> None:
        """Advance each truncation group t by counts[t] steps of dts[t] (groups in ascending truncation)."""
        d = np.ascontiguousarray(dts, dtype=np.float64)
        c = np.ascontiguousarray(counts, dtype=np.int64)
        fn = "lmw_ensemble_steps_graph" if graph else "lmw_ensemble_steps"
        L.call(fn, self._h, L.dptr(d), c.ctypes.data_as(L._i64p))


    def macro_steps(self, dts, substeps, count: int) -> None:
        """`count` macro steps; in each, group t advances substeps[t] steps of dts[t] (one HIP graph per macro step,
        the groups concurrent on their streams)."""
        d = np.ascontiguousarray(dts, dtype=np.float64)
        c = np.ascontiguousarray(substeps, dtype=np.int32)
        L.call("lmw_ensemble_macro_steps", self._h, L.dptr(d), c.ctypes.data_as(L._ip), int(count))

    def set_persistent(self, on: bool) -> None:
        """Macro steps through persistent multi-group Legendre kernels (mixed precision, at most 4 groups)."""
        L.call("lmw_ensemble_set_persistent", self._h, int(bool(on)))

    def autotune(self, force: bool = False):
        """Choose the Legendre tile shapes of every group (timed, or from the cache); returns (tiles (groups, 4),
        best_ms (groups, 2))."""
        ng = len(self.group_truncations)
        tiles = np.zeros(4 * ng, dtype=np.int32)
        ms = np.zeros(2 * ng)
        L.call("lmw_ensemble_autotune", self._h, int(bool(force)), tiles.ctypes.data_as(L._ip), L.dptr(ms))
        return tiles.reshape(ng, 4), ms.reshape(ng, 2)

    def set_tiles(self, group: int, tiles) -> None:
        t = np.ascontiguousarray(tiles, dtype=np.int32)
        L.call("lmw_ensemble_set_tiles", self._h, int(group), t.ctypes.data_as(L._ip))

    def steps_timed(self, dts, counts) -> np.ndarray:
        """Eager steps with HIP events around every kernel; returns (groups, 3) total milliseconds of the inverse
        Legendre, grid-point and forward Legendre kernels."""
        d = np.ascontiguousarray(dts, dtype=np.float64)
        c = np.ascontiguousarray(counts, dtype=np.int64)
        ms = np.zeros(3 * len(self.group_truncations))
        L.call("lmw_ensemble_steps_timed", self._h, L.dptr(d), c.ctypes.data_as(L._i64p), L.dptr(ms))
        return ms.reshape(-1, 3)


def fft(data, sign: int) -> np.ndarray:
    """Reference fft (cpp/sphere.cpp:21-45) on the GPU for a batch: data has shape (count, n) complex."""
    a = L.as_c128(data).copy()
    if a.ndim == 1:
        a = a[None, :]
    L.call("lmw_fft", a.ctypes.data_as(L._dp), a.shape[1], a.shape[0], sign)
    return a


def gauss_legendre(n: int):
    """Reference gauss_legendre (cpp/sphere.cpp:47-64) computed on the GPU."""
    x, w = np.empty(n), np.empty(n)
    L.call("lmw_gauss_legendre", n, L.dptr(x), L.dptr(w))
    return x, w


def tile_variant(kind: int, index: int):
    """Shape of a tile variant: kind 0 FP64 inverse members per block, 1 mixed inverse row chunk, 2 mixed forward
    (latitude chunk, waves), 3 FP64 forward (members per block of the FMA kernel, 0) or (0, waves of the MFMA
    kernel)."""
    a, b = C.c_int(), C.c_int()
    L.call("lmw_tile_variants", int(kind), int(index), C.byref(a), C.byref(b))
    return a.value, b.value


def device_info():
    name = C.create_string_buffer(256)
    cus, used = C.c_int(), C.c_size_t()
    L.call("lmw_device_info", name, 256, C.byref(cus), C.byref(used))
    return name.value.decode(), cus.value, used.value
