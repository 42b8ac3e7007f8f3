"""Host mirror of lmweather::Plane (reference cpp/plane.hpp:20-68) over the gfx950 C ABI, for a batch of members.

Member 0 is the one the reference-style accessors address; ``batch`` members share the parameters and step together.
Errors are ValueError / RuntimeError with the reference's messages.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import _lib as L


class Plane:
    def __init__(self, params: L.PlaneParams, batch: int = 1, precision: int = L.FP64):
        self.params = params
        h = C.c_void_p()
        L.call("lmw_plane_create", C.byref(params), batch, precision, C.byref(h))
        self._h = h
        self.batch, self.precision = batch, precision
        self.nx, self.ny = params.nx, params.ny

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                L.lib().lmw_plane_destroy(h)
            except Exception:
                pass
            self._h = None

    # geometry (plane.hpp:26-31)
    def dx(self) -> float:
        return self.params.lx / self.params.nx

    def dy(self) -> float:
        return self.params.ly / self.params.ny

    def x_center(self, i: int) -> float:
        return (i + 0.5) * self.dx()

    def y_center(self, j: int) -> float:
        return (j + 0.5) * self.dy()

    def set_params(self, **kw) -> None:
        for k, v in kw.items():
            setattr(self.params, k, v)
        L.call("lmw_plane_set_params", self._h, C.byref(self.params))

    # state
    def reset(self) -> None:
        L.call("lmw_plane_reset", self._h)

    def perturb(self, amplitude: float, seed: int, member: int = 0) -> None:
        L.call("lmw_plane_perturb", self._h, member, amplitude, seed)

    def fields(self, h: str, u: str, v: str, member: int = 0) -> None:
        """The Plane.fields operator (main.cpp:410-431): h = H + h(x, y) at the cell centres, u at the west faces, v
        at the interior south faces, from expressions in x, y and lx, ly, H, g, f0, beta, evaluated on the device."""
        L.call("lmw_plane_fields", self._h, member, L.text(h), L.text(u), L.text(v))

    def _rows(self, field):
        return self.ny + 1 if field == L.PLANE_V else self.ny

    def set_field(self, field: int, g, member: int = 0) -> None:
        g = L.as_f64(g).ravel()
        L.call("lmw_plane_set_field", self._h, member, field, L.dptr(g), g.size)

    def get_field(self, field: int, member: int = 0) -> np.ndarray:
        g = np.empty(self._rows(field) * self.nx)
        L.call("lmw_plane_get_field", self._h, member, field, L.dptr(g), g.size)
        return g.reshape(self._rows(field), self.nx)

    def set_h_from_grid(self, g, member: int = 0):
        self.set_field(L.PLANE_H, g, member)

    def set_u_from_grid(self, g, member: int = 0):
        self.set_field(L.PLANE_U, g, member)

    def set_v_from_grid(self, g, member: int = 0):
        self.set_field(L.PLANE_V, g, member)

    def h(self, member: int = 0):
        return self.get_field(L.PLANE_H, member)

    def u(self, member: int = 0):
        return self.get_field(L.PLANE_U, member)

    def v(self, member: int = 0):
        return self.get_field(L.PLANE_V, member)

    # stepping
    def step(self, dt: float) -> None:
        L.call("lmw_plane_step", self._h, dt)

    def set_tile(self, tile: int) -> None:
        """Step kernel: the row march (one launch per RK4 step) with rows per wavefront chosen from the occupancy
        (0, the default) or 1 = 32, 2 = 64, 3 = 128, 4 = 256 rows; or 5 = four stage launches. All bit-identical."""
        L.call("lmw_plane_set_tile", self._h, int(tile))

    def steps(self, dt: float, count: int) -> None:
        L.call("lmw_plane_steps", self._h, dt, count)

    def steps_timed(self, dt: float, count: int) -> float:
        """steps(dt, count), returning the time of the step kernels alone in ms (HIP events on the plane's stream)."""
        ms = C.c_double()
        L.call("lmw_plane_steps_timed", self._h, dt, count, C.byref(ms))
        return ms.value

    def run(self, seconds: float, dt: float) -> None:
        L.call("lmw_plane_run", self._h, seconds, dt)

    def stable_dt(self, safety: float = 0.5) -> float:
        out = C.c_double()
        L.call("lmw_plane_stable_dt", self._h, safety, C.byref(out))
        return out.value

    # diagnostics
    def _diag(self, which, member=0) -> float:
        out = C.c_double()
        L.call("lmw_plane_diag", self._h, member, which, C.byref(out))
        return out.value

    def time(self) -> float:
        return self._diag(L.PDIAG_TIME)

    def mean_depth(self, member: int = 0) -> float:
        return self._diag(L.PDIAG_MEAN_DEPTH, member)

    def energy(self, member: int = 0) -> float:
        return self._diag(L.PDIAG_ENERGY, member)

    def max_speed(self, member: int = 0) -> float:
        return self._diag(L.PDIAG_MAX_SPEED, member)

    def potential_vorticity_anomaly_rms(self, member: int = 0) -> float:
        return self._diag(L.PDIAG_PV_ANOMALY_RMS, member)

    def wave_speed(self) -> float:  # main.cpp:445
        return math.sqrt(self.params.gravity * self.params.depth)

    def rossby_radius(self) -> float:  # main.cpp:446-449
        if self.params.f0 == 0:
            raise ValueError("Plane.rossby_radius: f0 is zero")
        return math.sqrt(self.params.gravity * self.params.depth) / abs(self.params.f0)

    # zonal harmonics (plane.cpp:264-310), on the device
    def _harm(self, name, *args) -> float:
        out = C.c_double()
        L.call(name, self._h, *args, C.byref(out))
        return out.value

    def zonal_amplitude(self, m: int, j: int, member: int = 0) -> float:
        return self._harm("lmw_plane_zonal_amplitude", member, m, j)

    def zonal_crest(self, m: int, j: int, member: int = 0) -> float:
        return self._harm("lmw_plane_zonal_crest", member, m, j)

    def v_zonal_crest(self, m: int, j: int, member: int = 0) -> float:
        return self._harm("lmw_plane_v_zonal_crest", member, m, j)

    def harmonic_energy(self, m: int, member: int = 0) -> float:
        return self._harm("lmw_plane_harmonic_energy", member, m)

    # text output (main.cpp:464-493)
    def save_csv(self, path: str, member: int = 0) -> None:
        """Rows "x y h u v zeta" at the cell centres ("%.10g"), a blank line after each row."""
        L.call("lmw_plane_save_csv", self._h, member, os.fsencode(path))

    def save_row(self, row: int, path: str, member: int = 0) -> None:
        L.call("lmw_plane_save_row", self._h, member, row, os.fsencode(path))

    def save_column(self, column: int, path: str, member: int = 0) -> None:
        L.call("lmw_plane_save_column", self._h, member, column, os.fsencode(path))

    def tendency(self, member: int = 0):
        dh, du = np.empty(self.nx * self.ny), np.empty(self.nx * self.ny)
        dv = np.empty(self.nx * (self.ny + 1))
        L.call("lmw_plane_tendency", self._h, member, L.dptr(dh), L.dptr(du), L.dptr(dv))
        return dh.reshape(self.ny, self.nx), du.reshape(self.ny, self.nx), dv.reshape(self.ny + 1, self.nx)

    # point evaluation (bilinear on each field's own points, plane.cpp:245-263), on the host from the fields
    def _interp(self, x, y, kind, member=0):
        p = self.params
        dxx, dyy = p.lx / p.nx, p.ly / p.ny
        xs, ys = (0.0 if kind == 1 else 0.5), (0.0 if kind == 2 else 0.5)
        fx, fy = x / dxx - xs, y / dyy - ys
        rows = p.ny + 1 if kind == 2 else p.ny
        fl = math.floor(fx)
        i0, j0 = int(fl), int(math.floor(fy))
        ax, ay = fx - fl, fy - math.floor(fy)
        f = self.get_field(kind, member)

        def get(i, j):
            return f[max(0, min(rows - 1, j)), i % p.nx]

        return (1 - ax) * ((1 - ay) * get(i0, j0) + ay * get(i0, j0 + 1)) + ax * ((1 - ay) * get(i0 + 1, j0) + ay * get(i0 + 1, j0 + 1))

    def h_at(self, x, y, member=0):
        return self._interp(x, y, 0, member)

    def u_at(self, x, y, member=0):
        return self._interp(x, y, 1, member)

    def v_at(self, x, y, member=0):
        return self._interp(x, y, 2, member)
