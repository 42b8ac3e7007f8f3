"""Pinning of the oracle's text writers (oracle/writers.py), on the CPU.

1. The number format against the C library's own snprintf("%.10g") (the call the reference makes, cpp/main.cpp:84)
   over values of every magnitude and the rounding edge cases.
2. The layouts against the reference's own checks of its files (test.sh:154-163): 8 columns and nlat x nlon rows for
   the sphere CSV with a blank line after each latitude row, 6 columns for the plane CSV, "n ke pe" lines for
   n = 1..T, and Parseval: the kinetic spectrum of solid-body rotation u = 15 cos(lat) sums to 15^2 / 3 = 75
   (tests/sphere_spectrum.lms).
3. Each column against what it must hold: longitudes 360 k / nlon, latitudes asin(mu) in degrees, the centre
   positions of the plane, zonal means of a field that is constant along each row.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import writers as W

A, OM, G = 6.37122e6, 7.292e-5, 9.80616


def c_g(x: float) -> str:
    libc = ctypes.CDLL(None)
    buf = ctypes.create_string_buffer(64)
    libc.snprintf(buf, ctypes.c_size_t(64), b"%.10g", ctypes.c_double(x))
    return buf.value.decode()


def test_number_format_matches_c_snprintf():
    rng = np.random.default_rng(0)
    vals = [0.0, -0.0, 1.0, -1.0, 0.5, 1e-5, 1e-4, 123456789.0, 1234567890.0, 12345678901.0, 9999999999.5, 0.99999999995,
            1e300, -1e-300, 5e-324, 2.5, 1.00000000005, 6.02214076e23, 3.141592653589793, 2000.0, 15.0 * math.cos(0.3)]
    vals += list(rng.standard_normal(500) * 10.0 ** rng.integers(-12, 12, 500))
    vals += [float(np.nextafter(x, np.inf)) for x in vals[:20]]
    for x in vals:
        assert W.g(x) == c_g(x), x


def _fields(s, h, u, v):
    mu = s.gauss()[0]
    lat = np.arcsin(mu)[:, None] * np.ones((1, s.nlon))
    lon = 2 * np.pi * np.arange(s.nlon)[None, :] / s.nlon * np.ones((s.nlat, 1))
    s.set_state_grid(h(lat, lon), u(lat, lon), v(lat, lon))
    return mu


@pytest.fixture(scope="module")
def solid(oracle):
    s = oracle.OracleSphere(21, A, OM, G)
    mu = _fields(s, lambda la, lo: 2000.0 + 0 * la, lambda la, lo: 15.0 * np.cos(la), lambda la, lo: 0 * la)
    return s, mu


def test_sphere_spectrum_layout_and_parseval(solid):
    s, _ = solid
    ke, pe = s.spectrum()
    text = W.sphere_spectrum(ke, pe)
    lines = [l for l in text.split("\n") if l]
    assert len(lines) == 21 and {len(l.split()) for l in lines} == {3}
    assert [int(l.split()[0]) for l in lines] == list(range(1, 22))
    assert abs(sum(float(l.split()[1]) for l in lines) - 75.0) <= 1e-6  # test.sh:162
    assert text.endswith("\n") and "\n\n" not in text


def test_sphere_csv_layout(solid):
    s, mu = solid
    h, u, v, z = s.grid_height(), s.grid_u(), s.grid_v(), s.grid_vorticity()
    zero = np.zeros_like(h)
    text = W.sphere_csv(mu, h, u, v, z, zero, zero)
    rows = text.split("\n\n")
    assert len(rows) == 33 and rows[-1] == ""  # 32 latitude rows, each followed by a blank line
    lines = [l for l in text.split("\n") if l]
    assert len(lines) == 2048 and {len(l.split()) for l in lines} == {8}  # test.sh:156-158
    cols = np.array([[float(x) for x in l.split()] for l in lines]).reshape(32, 64, 8)
    assert np.allclose(cols[0, :, 0], 360.0 * np.arange(64) / 64, rtol=0, atol=1e-8)
    assert np.allclose(cols[:, 0, 1], np.degrees(np.arcsin(mu)), rtol=1e-9, atol=0)
    assert np.all(np.diff(cols[:, 0, 1]) > 0) and abs(cols[0, 0, 1] + cols[-1, 0, 1]) < 1e-8
    assert np.allclose(cols[:, :, 2], h, rtol=1e-9) and np.allclose(cols[:, :, 3], u, rtol=1e-9, atol=1e-9)
    assert np.all(cols[:, :, 6:] == 0)


def test_sphere_zonal_mean(solid):
    s, mu = solid
    h, u = s.grid_height(), s.grid_u()
    text = W.sphere_zonal_mean(mu, h, u)
    lines = [l.split() for l in text.split("\n") if l]
    assert len(lines) == 32 and {len(l) for l in lines} == {3}
    got = np.array([[float(x) for x in l] for l in lines])
    assert np.allclose(got[:, 1], 2000.0, rtol=1e-9)
    assert np.allclose(got[:, 2], 15.0 * np.cos(np.arcsin(mu)), rtol=1e-8)


def test_plane_writers_layout(oracle):
    # the plane of tests/determinism.lms, 40 x 20
    o = oracle.OraclePlane(40, 20, 4.0e6, 2.0e6, 9.81, 500.0, 1e-4, 2e-11)
    o.perturb(1.0, 99)
    for _ in range(5):
        o.step(600.0)
    h, u, v = o.get_field(0), o.get_field(1), o.get_field(2)
    text = W.plane_csv(40, 20, 4.0e6, 2.0e6, h, u, v)
    lines = [l for l in text.split("\n") if l]
    assert len(lines) == 800 and {len(l.split()) for l in lines} == {6}  # test.sh:157
    assert text.count("\n\n") == 20
    cols = np.array([[float(x) for x in l.split()] for l in lines]).reshape(20, 40, 6)
    assert np.allclose(cols[0, :, 0], (np.arange(40) + 0.5) * 1e5) and np.allclose(cols[:, 0, 1], (np.arange(20) + 0.5) * 1e5)
    assert np.allclose(cols[:, :, 2], h, rtol=1e-9)
    # the corner vorticity averaged to the centre, from the oracle's own zeta_at_corner
    zc = np.array([[0.25 * (o.zeta_at_corner(i, j) + o.zeta_at_corner(i + 1, j) + o.zeta_at_corner(i, j + 1) + o.zeta_at_corner(i + 1, j + 1))
                    for i in range(40)] for j in range(20)])
    assert np.allclose(cols[:, :, 5], zc, rtol=1e-9, atol=1e-18)
    row = [l.split() for l in W.plane_row(40, 20, 4.0e6, 2.0e6, h, 7).split("\n") if l]
    assert len(row) == 40 and np.allclose([float(r[1]) for r in row], h[7], rtol=1e-9)
    col = [l.split() for l in W.plane_column(40, 20, 4.0e6, 2.0e6, h, 3).split("\n") if l]
    assert len(col) == 20 and np.allclose([float(c[1]) for c in col], h[:, 3], rtol=1e-9)
