"""GPU initial-condition path, through the C ABI: the device-side Williamson set-ups (lmw_sphere_setup,
lmw_ensemble_setup; the Sphere.williamson1/2/5/6 operators of cpp/main.cpp:174-291) against the oracle set up from
the same formulas (oracle/setups.py) and against the reference's own known answers, and the batched ensemble set-up
(lmw_ensemble_set_state_grids) against the one-member path.
"""
import math

import numpy as np
import pytest

from oracle import setups as SU
from tests.helpers import max_rel

pytestmark = pytest.mark.gpu

A, OM, G = 6.37122e6, 7.292e-5, 9.80616
DEG = math.pi / 180.0
U0 = 2.0 * math.pi * A / (12.0 * 86400.0)


@pytest.fixture(scope="module")
def lmw():
    import lmweather_amd
    return lmweather_amd


def sphere(lmw, N):
    return lmw.Sphere(N, lmw.SphereParams(A, OM, G))


@pytest.mark.parametrize("which,alpha", [(SU.W1, math.pi / 2 - 0.05), (SU.W2, 0.0), (SU.W2, 0.7), (SU.W5, 0.0), (SU.W6, 0.0)])
def test_setup_matches_oracle(lmw, oracle, which, alpha):
    s = sphere(lmw, 42)
    s.setup(which, alpha)
    o = oracle.OracleSphere(42, A, OM, G)
    f = SU.apply(o, which, alpha, A, OM, G)
    for field in range(3):
        assert max_rel(s.get_spectral(field), o.get_spectral(field), scale=np.max(np.abs(o.get_spectral(0 if field == 1 else field)))) <= 1e-12
    if "hs" in f:
        assert max_rel(s.grid_orography(), f["hs"]) <= 1e-13
    if "q" in f:
        assert max_rel(s.get_spectral(3), o.get_spectral(3)) <= 1e-12
        assert s.advect_only() and s.has_tracer()
    # the tilt (Coriolis) and orography enter the tendencies; test case 2 is steady, so its tendencies are rounding
    # noise: compare on the scale of the balanced terms, 2 Omega times the state
    st = [o.get_spectral(f) for f in range(3)]
    scales = [2 * OM * np.max(np.abs(st[0])), 2 * OM * np.max(np.abs(st[0])), 2 * OM * np.max(np.abs(st[2]))]
    for a, b, sc in zip(s.tendencies(), o.tendencies(), scales):
        assert np.max(np.abs(a - b)) <= 1e-11 * sc
    assert abs(s.mean_height() - o.mean_height()) <= 1e-12 * o.mean_height()
    assert s.height_error_l2() == 0.0


def test_setup_known_answers(lmw):
    """sphere_williamson2.lms:22-30, sphere_orography.lms:23-27, sphere_tracer.lms:8-16 on the device set-ups."""
    s = sphere(lmw, 42)
    s.setup(SU.W2, 0.0)
    h0 = 29400.0 / G
    assert abs(s.max_wind() - U0) <= 0.02 and abs(s.mean_height() - 2363.02) <= 0.01
    assert abs(s.height_at(0.0, 0.0) - h0) <= 0.001
    s.setup(SU.W5, 0.0)
    surf = 5960.0 - (6371220.0 * 0.00007292 * 20.0 + 200.0) * 0.25 / 9.80616
    assert abs(s.max_wind() - 20.0) <= 0.02 and abs(s.mean_height() - 5619.9) <= 0.5
    assert abs(s.height_at(30.0 * DEG, 90.0 * DEG) - surf) <= 0.5
    t = sphere(lmw, 42)
    t.setup(SU.W1, math.pi / 2 - 0.05)
    c0 = t.tracer_at(0.0, 270.0 * DEG)
    assert abs(c0 - 1000.0) <= 5.0 and -4.0 < t.tracer_min() < 0.0 and 985.0 < t.tracer_max() < 1000.0
    t.run(6 * 86400.0, 1800.0)  # half way round: the bell at the antipode (sphere_tracer.lms:17-21)
    assert abs(t.tracer_at(0.0, 90.0 * DEG) - c0) <= 0.1
    assert abs(t.tracer_error_l2() - math.sqrt(2.0)) <= 0.001


@pytest.mark.parametrize("tilt", [0.0, 0.05, math.pi / 2 - 0.05, math.pi / 2])
def test_williamson2_steady(lmw, tilt):
    """sphere_williamson2.lms:5-19 on the device set-up."""
    s = sphere(lmw, 21)
    s.setup(SU.W2, tilt)
    m0, e0 = s.mean_height(), s.energy()
    s.run(432000.0, 1800.0)
    assert s.height_error_l2() < 1e-6 and s.height_error_linf() < 1e-6
    assert abs(s.mean_height() - m0) <= 1e-9 and abs(s.energy() / e0 - 1.0) <= 1e-9


def test_setup_errors(lmw):
    s = sphere(lmw, 21)
    with pytest.raises(ValueError, match="^angle must be a finite number$"):
        s.setup(SU.W2, float("nan"))
    with pytest.raises(ValueError, match="unknown set-up"):
        s.setup(3, 0.0)
    e = lmw.Ensemble([21, 21], lmw.SphereParams(A, OM, G))
    with pytest.raises(ValueError, match="^a tracer needs a single-member sphere$"):
        e.setup(0, 2, SU.W1, 0.0)


def test_ensemble_setup_matches_sphere(lmw):
    """Members of two truncations, set up in batched launches, equal the one-member set-up."""
    truncs = [42, 42, 42, 85, 85, 42]  # groups: T42 members 0, 1, 2, 5; T85 members 3, 4
    e = lmw.Ensemble(truncs, lmw.SphereParams(A, OM, G))
    e.setup(0, 4, SU.W6, 0.0)  # a run of three T42 members, then one T85 member
    e.setup(4, 2, SU.W5, 0.0)
    for i, (N, which) in enumerate(zip(truncs, [SU.W6] * 4 + [SU.W5] * 2)):
        s = sphere(lmw, N)
        s.setup(which, 0.0)
        for field in range(3):
            assert max_rel(e.get_spectral(i, field), s.get_spectral(field), scale=np.max(np.abs(s.get_spectral(0)))) <= 1e-14, (i, field)
        assert np.array_equal(e.grid(i, 5), s.grid_orography())


@pytest.mark.parametrize("precision", [0, 3])
def test_set_state_grids_matches_one_member_path(lmw, precision):
    from lmweather_amd.initial import random_balanced_state
    N, count = 42, 37  # more than one chunk of 32
    base = sphere(lmw, N)
    states = [random_balanced_state(base, 12, 100 + i) for i in range(count)]
    h = np.stack([x[0] for x in states])
    u = np.stack([x[1] for x in states])
    v = np.stack([x[2] for x in states])
    a = lmw.Ensemble([N] * (count + 2), lmw.SphereParams(A, OM, G), precision)
    b = lmw.Ensemble([N] * (count + 2), lmw.SphereParams(A, OM, G), precision)
    a.set_state_grids(1, count, h, u, v)
    for i in range(count):
        b.set_state_grid(1 + i, h[i], u[i], v[i])
    for i in range(1, count + 1):
        for field in range(3):
            assert np.array_equal(a.get_spectral(i, field), b.get_spectral(i, field)), (i, field)
        assert a.diag(i, 6) == 0.0  # error_l2 against the reference marked by the set-up
    assert np.all(a.get_spectral(0, 2) == 0) and np.all(a.get_spectral(count + 1, 2) == 0)
    with pytest.raises(ValueError, match="^grid size mismatch$"):
        a.set_state_grids(0, 2, h[:1], u[:1], v[:1])
    with pytest.raises(ValueError, match="^member index out of range$"):
        a.set_state_grids(count, 3, h[:3], u[:3], v[:3])


def test_orography_planes_shared_and_private(lmw):
    """Members with bit-identical orography share one device plane; setting a member's orography never changes
    another member's, whichever plane they shared (each member's orography, energy and tendencies equal those of a
    one-member sphere given the same state and orography)."""
    from lmweather_amd.initial import gaussian_orography, random_balanced_state
    N = 21
    base = sphere(lmw, N)
    h, u, v = random_balanced_state(base, 8, 3)
    ta, tb, tc = (gaussian_orography(base, 4, seed) for seed in (5, 6, 7))
    e = lmw.Ensemble([N] * 4, lmw.SphereParams(A, OM, G))
    for i in range(4):
        e.set_state_grid(i, h, u, v)
    plan = [ta, ta, ta, None]  # member 3 stays flat
    for i, t in enumerate(plan):
        if t is not None:
            e.set_orography_grid(i, t)
    e.set_orography_grid(1, tb)  # leaves a shared plane
    e.set_orography_grid(0, tb)  # joins member 1's plane
    e.set_orography_grid(2, tc)  # sole user of the first plane: rewritten in place
    e.set_orography_grid(1, ta)
    plan = [tb, ta, tc, None]
    e.setup(3, 1, SU.W5, 0.0)  # the cone: a plane of its own
    e.set_orography_grid(0, tc)  # shares member 2's plane
    plan = [tc, ta, tc, "cone"]
    cone = None
    for i, t in enumerate(plan):
        s = sphere(lmw, N)
        if isinstance(t, str):
            s.setup(SU.W5, 0.0)
            cone = s
        else:
            s.set_orography_grid(t)
            s.set_state_grid(h, u, v)
        assert np.array_equal(e.grid(i, 5), s.grid_orography()), i
        for field in range(3):
            assert np.array_equal(e.get_spectral(i, field), s.get_spectral(field)), (i, field)
        assert e.diag(i, 3) == s.potential_energy(), i  # potential energy reads the orography
    assert cone is not None
