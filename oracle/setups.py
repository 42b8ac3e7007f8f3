"""The analytic initial states of Williamson et al. (1992, J. Comput. Phys. 102:211-224) that the reference's
Sphere.williamson1/2/5/6 operators set up (cpp/main.cpp:174-291), restated with numpy on a model's Gauss grid.

TEST INFRASTRUCTURE ONLY: the tests set the oracle sphere up with these grids and compare the library's device-side
set-up (lmw_sphere_setup) with it.  lat = asin(mu_j), lon = 2 pi k / nlon (sphere.cpp:120-121).
"""
from __future__ import annotations

import math

import numpy as np

W1, W2, W5, W6 = 1, 2, 5, 6


def grids(mu, nlon, which, alpha=0.0, radius=6.37122e6, omega=7.292e-5, gravity=9.80616):
    """dict with h, u, v (nlat x nlon), the axis tilt, and hs (W5) or q (W1) where the case sets them."""
    pi = math.pi
    lat = np.arcsin(np.asarray(mu, dtype=np.float64))[:, None] * np.ones((1, nlon))
    lon = (2 * pi * np.arange(nlon) / nlon)[None, :] * np.ones((len(mu), 1))
    a, om, g = radius, omega, gravity
    cl, sl = np.cos(lat), np.sin(lat)
    out = {"tilt": 0.0}
    if which in (W1, W2):  # main.cpp:174-191, :263-291
        u0 = 2 * pi * a / (12 * 86400.0)
        out["tilt"] = alpha
        out["u"] = u0 * (cl * math.cos(alpha) + sl * np.cos(lon) * math.sin(alpha))
        out["v"] = -u0 * np.sin(lon) * math.sin(alpha)
        if which == W2:
            t = -np.cos(lon) * cl * math.sin(alpha) + sl * math.cos(alpha)
            out["h"] = 2.94e4 / g - (a * om * u0 + 0.5 * u0 * u0) * t * t / g
        else:
            out["h"] = np.full_like(lat, 1000.0)
            R, lonc, latc = a / 3, 1.5 * pi, 0.0
            c = math.sin(latc) * sl + math.cos(latc) * cl * np.cos(lon - lonc)
            rr = a * np.arccos(np.clip(c, -1.0, 1.0))
            out["q"] = np.where(rr < R, 500.0 * (1 + np.cos(pi * rr / R)), 0.0)
    elif which == W5:  # main.cpp:230-252
        u0, h0, hs0, R, lonc, latc = 20.0, 5960.0, 2000.0, pi / 9, 1.5 * pi, pi / 6
        rr = np.sqrt(np.minimum(R * R, (lon - lonc) ** 2 + (lat - latc) ** 2))
        hs = hs0 * (1 - rr / R)
        out["hs"] = hs
        out["u"] = u0 * cl
        out["v"] = np.zeros_like(lat)
        out["h"] = h0 - (a * om * u0 + 0.5 * u0 * u0) * sl * sl / g - hs
    elif which == W6:  # main.cpp:192-214
        w, K, R, h0 = 7.848e-6, 7.848e-6, 4.0, 8000.0
        out["u"] = a * w * cl + a * K * cl ** (R - 1) * (R * sl * sl - cl * cl) * np.cos(R * lon)
        out["v"] = -a * K * R * cl ** (R - 1) * sl * np.sin(R * lon)
        A = w / 2 * (2 * om + w) * cl * cl + K * K / 4 * cl ** (2 * R) * ((R + 1) * cl * cl + (2 * R * R - R - 2) - 2 * R * R / (cl * cl))
        B = 2 * (om + w) * K / ((R + 1) * (R + 2)) * cl ** R * ((R * R + 2 * R + 2) - (R + 1) ** 2 * cl * cl)
        C = K * K / 4 * cl ** (2 * R) * ((R + 1) * cl * cl - (R + 2))
        out["h"] = h0 + a * a / g * (A + B * np.cos(R * lon) + C * np.cos(2 * R * lon))
    else:
        raise ValueError(f"unknown set-up {which}")
    return out


def apply(sphere, which, alpha=0.0, radius=6.37122e6, omega=7.292e-5, gravity=9.80616):
    """Set an oracle.pyapi.OracleSphere up as the reference operator does: tilt, orography, state, tracer."""
    mu = sphere.gauss()[0]
    f = grids(mu, sphere.nlon, which, alpha, radius, omega, gravity)
    sphere.set_axis_tilt(f["tilt"])
    if "hs" in f:
        sphere.set_orography_grid(f["hs"])
    sphere.set_state_grid(f["h"], f["u"], f["v"])
    if "q" in f:
        sphere.set_tracer_grid(f["q"])
        sphere.set_advect_only(True)
    return f
