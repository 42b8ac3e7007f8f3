"""Pinning of the oracle's Williamson set-ups (oracle/setups.py), on the CPU, by the reference's own known answers
for them (tests/sphere_williamson2.lms, sphere_orography.lms, sphere_tracer.lms) and by closed forms: test case 2
is a steady state for every axis tilt, test case 6's vorticity is 2w sin(lat) - K sin(lat) cos^R(lat) (R^2 + 3R + 2)
cos(R lon) with zero divergence, test case 5's free surface h + hs is test case 2's balanced surface.
"""
import math

import numpy as np
import pytest

from oracle import setups as SU

A, OM, G = 6.37122e6, 7.292e-5, 9.80616
DEG = math.pi / 180.0
U0 = 2.0 * math.pi * A / (12.0 * 86400.0)


def make(oracle, N, which, alpha=0.0):
    s = oracle.OracleSphere(N, A, OM, G)
    f = SU.apply(s, which, alpha, A, OM, G)
    return s, f


def test_williamson2_known_values(oracle):
    """sphere_williamson2.lms:22-30 at T42."""
    s, _ = make(oracle, 42, SU.W2)
    h0 = 29400.0 / G
    assert abs(s.max_wind() - U0) <= 0.02
    assert abs(s.height_at(0.0, 0.0) - h0) <= 0.001
    assert abs(s.height_at(90.0 * DEG, 0.0) - (h0 - (A * 0.00007292 * U0 + 0.5 * U0 * U0) / G)) <= 0.001
    assert abs(s.mean_height() - 2363.02) <= 0.01


@pytest.mark.parametrize("tilt", [0.0, 0.05, math.pi / 2 - 0.05, math.pi / 2])
def test_williamson2_steady_for_every_tilt(oracle, tilt):
    """sphere_williamson2.lms:5-19: T21, five days of 1800 s steps leave the state unchanged to 1e-6."""
    s, _ = make(oracle, 21, SU.W2, tilt)
    m0, e0 = s.mean_height(), s.energy()
    s.run(432000.0, 1800.0)
    l2, linf = s.height_errors()
    assert l2 < 1e-6 and linf < 1e-6
    assert abs(s.mean_height() - m0) <= 1e-9 and abs(s.energy() / e0 - 1.0) <= 1e-9


def test_williamson2_wrong_tilt_is_not_steady(oracle):
    """sphere_williamson2.lms:31-34: the tilted flow on an untilted planet moves."""
    s = oracle.OracleSphere(21, A, OM, G)
    f = SU.grids(s.gauss()[0], s.nlon, SU.W2, 1.5, A, OM, G)
    s.set_state_grid(f["h"], f["u"], f["v"])  # axis tilt left at 0
    s.run(86400.0, 1800.0)
    assert s.height_errors()[0] > 0.001


def test_williamson5_known_values(oracle):
    """sphere_orography.lms:23-27 at T42."""
    s, f = make(oracle, 42, SU.W5)
    surf = 5960.0 - (6371220.0 * 0.00007292 * 20.0 + 200.0) * 0.25 / 9.80616
    assert abs(s.max_wind() - 20.0) <= 0.02
    assert abs(s.mean_height() - 5619.9) <= 0.5
    assert abs(s.height_at(30.0 * DEG, 270.0 * DEG) + 2000.0 - surf) <= 200.0
    assert abs(s.height_at(30.0 * DEG, 90.0 * DEG) - surf) <= 0.5
    # the free surface is test case 2's balanced surface for u0 = 20 m/s; the cone peaks at 2000 m
    lat = np.arcsin(s.gauss()[0])[:, None]
    want = 5960.0 - (A * OM * 20.0 + 200.0) * np.sin(lat) ** 2 / G
    assert np.max(np.abs(f["h"] + f["hs"] - want)) <= 1e-9
    assert 1900.0 < f["hs"].max() <= 2000.0 and f["hs"].min() == 0.0


def test_williamson1_known_values(oracle):
    """sphere_tracer.lms:8-16 at T42 over the poles."""
    s, f = make(oracle, 42, SU.W1, math.pi / 2 - 0.05)
    q = s.grid_tracer()
    assert abs(s.tracer_at(0.0, 270.0 * DEG) - 1000.0) <= 5.0
    assert -4.0 < q.min() < 0.0 and 985.0 < q.max() < 1000.0
    assert abs(s.tracer_at(0.0, 90.0 * DEG)) <= 1.5
    assert s.tracer_errors()[0] == 0.0
    assert abs(s.max_wind() - U0) <= 0.02


def test_williamson6_vorticity_closed_form(oracle):
    s, _ = make(oracle, 21, SU.W6)
    lat = np.arcsin(s.gauss()[0])[:, None]
    lon = (2 * np.pi * np.arange(s.nlon) / s.nlon)[None, :]
    w = K = 7.848e-6
    R = 4
    zeta = 2 * w * np.sin(lat) - K * np.sin(lat) * np.cos(lat) ** R * (R * R + 3 * R + 2) * np.cos(R * lon)
    assert np.max(np.abs(s.grid_vorticity() - zeta)) <= 1e-10 * np.max(np.abs(zeta))
    d = s.get_spectral(1)
    assert np.max(np.abs(d)) <= 1e-12 * np.max(np.abs(s.get_spectral(0)))
