"""GPU passive tracer (reference cpp/sphere.cpp:236-299 and the tracer part of step, :407-436): parity with the CPU
oracle with and without frozen dynamics, and the closed-form checks of the reference's tests/sphere_tracer.lms
(Williamson test case 1: a cosine bell carried once around the globe in 12 days)."""
import math

import numpy as np
import pytest

from tests.helpers import OracleTransforms, rel_l2

pytestmark = pytest.mark.gpu

A_TC, OM, G = 6371220.0, 7.292e-5, 9.80616


@pytest.fixture(scope="module")
def lmw():
    import lmweather_amd
    return lmweather_amd


def williamson1_grids(lat, lon, alpha, radius):
    """Test case 1 set-up of the reference's Sphere.williamson1 (cpp/main.cpp:271-293): winds of a solid-body
    rotation about an axis tilted by alpha, h = 1000, a cosine bell of 1000 m and radius a/3 at (0 N, 270 E)."""
    lat, lon = lat[:, None], lon[None, :]
    u0 = 2 * math.pi * radius / (12 * 86400.0)
    R, lonc, latc = radius / 3, 1.5 * math.pi, 0.0
    u = u0 * (np.cos(lat) * math.cos(alpha) + np.sin(lat) * np.cos(lon) * math.sin(alpha))
    v = -u0 * np.sin(lon) * math.sin(alpha) * np.ones_like(lat)
    c = math.sin(latc) * np.sin(lat) + math.cos(latc) * np.cos(lat) * np.cos(lon - lonc)
    rr = radius * np.arccos(np.clip(c, -1.0, 1.0))
    q = np.where(rr < R, 500.0 * (1 + np.cos(math.pi * rr / R)), 0.0)
    return np.full_like(q, 1000.0), u, v, q


def tc1(model, alpha):
    model.set_axis_tilt(alpha)
    h, u, v, q = williamson1_grids(model.lats(), model.lons(), alpha, model.params.radius)
    model.set_state_grid(h, u, v)
    model.set_tracer_grid(q)
    model.set_advect_only(True)
    return model


def _oracle_grid_coords(o):
    mu, _ = o.gauss()
    return np.arcsin(mu), 2 * np.pi * np.arange(o.nlon) / o.nlon


def test_errors_without_tracer(lmw):
    s = lmw.Sphere(21, lmw.SphereParams(A_TC, OM, G))
    with pytest.raises(RuntimeError, match="no tracer: set one with tracer\\(\\) first"):
        s.grid_tracer()
    with pytest.raises(RuntimeError, match="no tracer"):
        s.tracer_mean()
    s.set_advect_only(True)
    with pytest.raises(ValueError, match="advect_only: set a tracer first \\(Sphere.tracer\\)"):
        s.run(3600.0, 600.0)
    assert s.has_tracer() is False and s.advect_only() is True


def test_advect_only_parity(lmw, oracle):
    """TC1 at T21 over the poles: 12 one-hour steps through run() against the oracle; the dynamics stay frozen."""
    N, alpha = 21, math.pi / 2 - 0.05
    s = lmw.Sphere(N, lmw.SphereParams(A_TC, OM, G))
    o = oracle.OracleSphere(N, A_TC, OM, G)
    lat, lon = _oracle_grid_coords(o)
    h, u, v, q = williamson1_grids(lat, lon, alpha, A_TC)
    for m in (s, o):
        m.set_axis_tilt(alpha)
        m.set_state_grid(h, u, v)
        m.set_tracer_grid(q)
        m.set_advect_only(True)
    z0 = s.get_spectral(0)
    assert rel_l2(s.get_spectral(3), o.get_spectral(3)) <= 1e-13
    s.run(12 * 3600.0, 3600.0)
    o.run(12 * 3600.0, 3600.0)
    e = rel_l2(s.get_spectral(3), o.get_spectral(3))
    print(f"T21 advect-only 12 h: rel-L2(q) vs oracle {e:.2e}")
    assert e <= 1e-12
    assert np.array_equal(s.get_spectral(0), z0)  # frozen dynamics
    assert rel_l2(s.grid_tracer(), o.grid_tracer()) <= 1e-12
    l2, linf = o.tracer_errors()
    assert abs(s.tracer_error_l2() - l2) <= 1e-10 * l2 and abs(s.tracer_error_linf() - linf) <= 1e-10 * linf
    assert abs(s.tracer_mean() - o.tracer_mean()) <= 1e-12 * abs(o.tracer_mean())
    assert abs(s.tracer_mass() - o.tracer_mass()) <= 1e-12 * abs(o.tracer_mass())
    assert abs(s.tracer_at(0.3, 1.0) - o.tracer_at(0.3, 1.0)) <= 1e-9
    assert s.time() == o.time()


def test_tracer_with_dynamics_and_diffusion_parity(lmw, oracle):
    """Full dynamics, a tracer 1 + sin(lat) + bell, del^4 diffusion: 70 steps through run() (one periodic check)."""
    from lmweather_amd.initial import random_balanced_state
    N, nu4 = 21, 1e16
    p = lmw.SphereParams(6.371e6, OM, G, nu4)
    s = lmw.Sphere(N, p)
    o = oracle.OracleSphere(N, 6.371e6, OM, G, nu4)
    h, u, v = random_balanced_state(OracleTransforms(o, p), 12, 3)
    lat, lon = _oracle_grid_coords(o)
    q = 1.0 + np.sin(lat)[:, None] + williamson1_grids(lat, lon, 0.0, 6.371e6)[3] / 1000.0
    for m in (s, o):
        m.set_state_grid(h, u, v)
        m.set_tracer_grid(q)
    s.run(70 * 1200.0, 1200.0)
    o.run(70 * 1200.0, 1200.0)
    for f in range(4):
        e = rel_l2(s.get_spectral(f), o.get_spectral(f))
        print(f"field {f}: rel-L2 vs oracle {e:.2e}")
        assert e <= 1e-12, f
    assert rel_l2(s.grid_height(), o.grid_height()) <= 1e-12
    assert abs(s.tracer_mass() - o.tracer_mass()) <= 1e-12 * abs(o.tracer_mass())


def test_tracer_spectral_set_and_graph_replay(lmw, oracle):
    """A tracer set from its coefficients; steps() (HIP graph replay) agrees with step()."""
    from tests.helpers import random_spec
    N = 21
    s1, s2 = (lmw.Sphere(N, lmw.SphereParams(6.371e6, OM, G)) for _ in range(2))
    o = oracle.OracleSphere(N, 6.371e6, OM, G)
    from lmweather_amd.initial import random_balanced_state
    h, u, v = random_balanced_state(OracleTransforms(o, lmw.SphereParams(6.371e6, OM, G)), 10, 5)
    qs = random_spec(N, 9, 1.0)
    for m in (s1, s2, o):
        m.set_state_grid(h, u, v)
        m.set_spectral(3, qs)
    assert s1.has_tracer()
    s1.steps(900.0, 8)
    for _ in range(8):
        s2.step(900.0)
        o.step(900.0)
    assert np.array_equal(s1.get_spectral(3), s2.get_spectral(3))
    assert rel_l2(s1.get_spectral(3), o.get_spectral(3)) <= 1e-12


def test_williamson1_closed_forms(lmw):
    """The checks of the reference's tests/sphere_tracer.lms for TC1 at T42 over the poles with 30-minute steps."""
    s = tc1(lmw.Sphere(42, lmw.SphereParams(A_TC, OM, G)), math.pi / 2 - 0.05)
    q0 = s.tracer_mean()
    c0 = s.tracer_at(0.0, math.radians(270.0))
    assert abs(c0 - 1000.0) <= 5.0
    assert -4.0 < s.tracer_min() < 0.0
    assert 985.0 < s.tracer_max() < 1000.0
    assert abs(s.tracer_at(0.0, math.radians(90.0))) <= 1.5
    assert s.tracer_error_l2() == 0.0
    assert abs(s.max_wind() - 2 * math.pi * A_TC / (12 * 86400.0)) <= 0.02
    s.run(6 * 86400.0, 1800.0)
    assert abs(s.tracer_at(0.0, math.radians(90.0)) - c0) <= 0.1
    assert abs(s.tracer_error_l2() - math.sqrt(2.0)) <= 0.001
    assert abs(s.tracer_error_linf() - 1.0) <= 0.002
    s.run(6 * 86400.0, 1800.0)
    print(f"TC1 T42 day 12: l2 {s.tracer_error_l2():.3e} linf {s.tracer_error_linf():.3e} min {s.tracer_min():.3f} max {s.tracer_max():.3f}")
    assert s.tracer_error_l2() < 0.001 and s.tracer_error_linf() < 0.001
    assert abs(s.tracer_at(0.0, math.radians(270.0)) - c0) <= 0.1
    assert abs(s.tracer_mean() - q0) <= 1e-9
    assert abs(s.time() - 12 * 86400.0) <= 0.001
    assert s.height_error_l2() == 0.0
    # T21 with one-hour steps
    c = tc1(lmw.Sphere(21, lmw.SphereParams(A_TC, OM, G)), 0.0)
    cc0 = c.tracer_at(0.0, math.radians(270.0))
    c.run(12 * 86400.0, 3600.0)
    assert c.tracer_error_l2() < 0.003
    assert abs(c.tracer_at(0.0, math.radians(270.0)) - cc0) <= 0.3
