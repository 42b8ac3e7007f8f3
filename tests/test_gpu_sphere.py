"""GPU parity of the sphere (reference cpp/sphere.cpp) against the CPU oracle, through the C ABI.

Tolerances: the transforms and one tendency evaluation agree to a few ulps of the largest value (different summation
order: equatorial folding, FFT twiddles rounded once, FMA contraction); multi-step states to 1e-12 relative L2 in h
(the bar for the 5-day T42 run); diagnostics to 1e-12 relative.
"""
import math

import numpy as np
import pytest

from tests.helpers import OracleTransforms, max_rel, random_spec, rel_l2

pytestmark = pytest.mark.gpu

A, OM, G = 6.371e6, 7.292e-5, 9.80616


@pytest.fixture(scope="module")
def lmw():
    import lmweather_amd
    return lmweather_amd


def make_pair(lmw, oracle, N, radius=A, omega=OM, gravity=G, nu4=0.0):
    p = lmw.SphereParams(radius, omega, gravity, nu4)
    return lmw.Sphere(N, p), oracle.OracleSphere(N, radius, omega, gravity, nu4), p


def balanced(lmw, oracle, N, nmax, seed):
    from lmweather_amd.initial import random_balanced_state
    s, o, p = make_pair(lmw, oracle, N)
    h, u, v = random_balanced_state(OracleTransforms(o, p), nmax, seed)
    return s, o, (h, u, v)


def test_gauss_legendre(lmw, oracle):
    for n in (8, 32, 64, 256):
        x, w = lmw.gauss_legendre(n)
        xr, wr = oracle.gauss_legendre(n)
        assert np.max(np.abs(x - xr)) <= 2e-16
        assert np.max(np.abs(w - wr)) <= 1e-15 * np.max(wr)


def test_fft_matches_oracle_and_dft(lmw, oracle):
    """lmw_fft runs the grid kernel's four-step FFT (fft_lds.h: fft2) for the lengths 8..256 that the time step uses
    and the radix-2 kernel beyond them; both against the oracle's fft (sphere.cpp:21-45) and a direct DFT."""
    rng = np.random.default_rng(1)
    for n in (8, 16, 32, 64, 128, 256, 512):
        a = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
        for sign in (-1, 1):
            got = lmw.fft(a, sign)
            dft = np.fft.fft(a, axis=1) if sign < 0 else n * np.fft.ifft(a, axis=1)  # unnormalized, e^(sign 2 pi i nk/N)
            assert np.max(np.abs(got - dft)) <= 1e-13 * n
            for r in range(3):
                ref = oracle.fft(a[r], sign)
                assert np.max(np.abs(got[r] - ref)) <= 1e-13 * n
    with pytest.raises(ValueError, match="fft length must be a power of two"):
        lmw.fft(np.zeros((1, 12)), -1)


@pytest.mark.parametrize("N", [21, 42, 170])
def test_tables(lmw, oracle, N):
    s, o, _ = make_pair(lmw, oracle, N)
    P, H = s.tables()
    Pr, Hr = o.tables()
    scale = np.max(np.abs(Pr))
    assert np.max(np.abs(P - Pr)) <= 1e-13 * scale
    assert np.max(np.abs(H - Hr)) <= 1e-13 * np.max(np.abs(Hr))
    mu, w = s.gauss_mu(), s.gauss_weights()
    mur, wr = o.gauss()
    assert np.max(np.abs(mu - mur)) <= 2e-16


@pytest.mark.parametrize("N", [21, 42, 85])
def test_transforms(lmw, oracle, N):
    s, o, _ = make_pair(lmw, oracle, N)
    spec = random_spec(N, 3)
    g = s.synthesize_scalar(spec)
    gr = o.synthesize(spec)
    assert max_rel(g, gr) <= 1e-13
    back = s.analyze_scalar(gr)
    assert max_rel(back, o.analyze(gr)) <= 1e-13
    assert max_rel(back, spec) <= 1e-12  # round trip (tests/core.cpp:60-71)
    rng = np.random.default_rng(4)
    Ag, Bg = rng.standard_normal(gr.shape), rng.standard_normal(gr.shape)
    assert max_rel(s.analyze_divergence(Ag, Bg), o.analyze_divergence(Ag, Bg)) <= 1e-12
    z, d = random_spec(N, 5, 1e-5), random_spec(N, 6, 1e-6)
    U, V = s.winds_spectral_to_grid(z, d)
    Ur, Vr = o.winds(z, d)
    assert max_rel(U, Ur) <= 1e-12 and max_rel(V, Vr) <= 1e-12


@pytest.mark.parametrize("N", [21, 42])
def test_state_and_tendencies(lmw, oracle, N):
    s, o, (h, u, v) = balanced(lmw, oracle, N, N // 2, 1)
    s.set_state_grid(h, u, v)
    o.set_state_grid(h, u, v)
    zscale = np.max(np.abs(o.get_spectral(0)))
    for f in range(3):  # delta is round-off only (the flow is non-divergent): measure it on the vorticity scale
        assert max_rel(s.get_spectral(f), o.get_spectral(f), zscale if f == 1 else None) <= 1e-12
    t = s.tendencies()
    tr = o.tendencies()
    for a, b in zip(t, tr):
        assert max_rel(a, b) <= 1e-11


def test_c1_five_days_t42(lmw, oracle):
    """Config C1: T42, a = 6.371e6, RK4 dt = 600 s for 5 days: rel-L2(h) <= 1e-12 against the oracle."""
    s, o, (h, u, v) = balanced(lmw, oracle, 42, 21, 1)
    s.set_state_grid(h, u, v)
    o.set_state_grid(h, u, v)
    e0, e0r = s.energy(), o.energy()
    z0, z0r = s.potential_enstrophy(), o.potential_enstrophy()
    m0r = o.mean_height()
    s.steps(600.0, 720)
    for _ in range(720):
        o.step(600.0)
    hg, hr = s.grid_height(), o.grid_height()
    err = rel_l2(hg, hr)
    print(f"C1 rel-L2(h) = {err:.3e}  u {rel_l2(s.grid_u(), o.grid_u()):.3e}")
    assert err <= 1e-12
    assert rel_l2(s.grid_u(), o.grid_u()) <= 1e-10
    # drift at least as good as the oracle's (to the oracle's own scale)
    de, der = abs(s.energy() - e0) / e0, abs(o.energy() - e0r) / e0r
    dz, dzr = abs(s.potential_enstrophy() - z0) / z0, abs(o.potential_enstrophy() - z0r) / z0r
    assert de <= der + 1e-13 and dz <= dzr + 1e-12
    assert abs(s.mean_height() - m0r) <= 1e-12 * m0r


def test_c2_t170_state_tendencies_and_steps(lmw, oracle):
    """Config C2's state (T170, balanced flow to n = 85 over 8 seeded mountains, dt = 150 s): set-up, tendencies and
    8 FP64 steps against the oracle; the BF16x6 Legendre mode's distance from FP64 after the same steps."""
    from lmweather_amd.initial import gaussian_orography
    s, o, (h, u, v) = balanced(lmw, oracle, 170, 85, 1)
    hs = gaussian_orography(OracleTransforms(o, s.params), 8, 2)
    m = lmw.Sphere(170, s.params)
    for x in (s, o, m):
        x.set_orography_grid(hs)
        x.set_state_grid(h - hs, u, v)
    m.set_precision(lmw.BF16X6)
    zscale = np.max(np.abs(o.get_spectral(0)))
    for f in range(3):
        assert max_rel(s.get_spectral(f), o.get_spectral(f), zscale if f == 1 else None) <= 1e-12
    for f, (a, b) in enumerate(zip(s.tendencies(), o.tendencies())):
        print(f"C2 T170 tendency {f}: rel-L2 {rel_l2(a, b):.2e}, max-rel {max_rel(a, b):.2e}")
        assert rel_l2(a, b) <= 1e-11 and max_rel(a, b) <= 1e-10
    s.steps(150.0, 8)
    m.steps(150.0, 8)
    for _ in range(8):
        o.step(150.0)
    err = rel_l2(s.grid_height(), o.grid_height())
    emx = rel_l2(m.grid_height(), o.grid_height())
    print(f"C2 T170, 8 steps: FP64 rel-L2(h) vs oracle {err:.2e}; BF16x6 {emx:.2e}")
    assert err <= 1e-12
    assert emx <= 1e-8
    assert m.mean_height() == s.mean_height()


def test_diagnostics_and_errors(lmw, oracle):
    s, o, (h, u, v) = balanced(lmw, oracle, 42, 21, 7)
    s.set_state_grid(h, u, v)
    o.set_state_grid(h, u, v)
    for _ in range(3):
        o.step(600.0)
    s.steps(600.0, 3)
    for name in ("energy", "kinetic_energy", "potential_energy", "potential_enstrophy", "max_wind", "mean_height"):
        assert abs(getattr(s, name)() - getattr(o, name)()) <= 1e-12 * abs(getattr(o, name)()), name
    assert abs(s.stable_dt() - o.stable_dt()) <= 1e-12 * o.stable_dt()
    l2r, linfr = o.height_errors()
    assert abs(s.height_error_l2() - l2r) <= 1e-9 * l2r
    ke, pe = s.spectrum()
    ker, per = o.spectrum()
    assert max_rel(ke, ker) <= 1e-12 and max_rel(pe, per) <= 1e-12
    for lat, lon in ((0.3, 1.0), (-1.2, 4.0)):
        assert abs(s.height_at(lat, lon) - o.height_at(lat, lon)) <= 1e-9
        assert abs(s.wind_at(lat, lon, True) - o.wind_at(lat, lon, True)) <= 1e-9
        assert abs(s.vorticity_at(lat, lon) - o.vorticity_at(lat, lon)) <= 1e-15
    assert abs(s.mode_coefficient(5, 2, True) - o.mode_coefficient(5, 2, True)) <= 1e-9
    # run() refusal text (sphere.cpp:463-466)
    with pytest.raises(ValueError) as e1:
        s.run(3600.0, 5000.0)
    with pytest.raises(ValueError) as e2:
        o.run(3600.0, 5000.0)
    assert str(e1.value).split("limit")[0] == str(e2.value).split("limit")[0]


def test_orography_tilt_diffusion_perturb(lmw, oracle):
    from lmweather_amd.initial import gaussian_orography
    s, o, (h, u, v) = balanced(lmw, oracle, 42, 21, 11)
    p = lmw.SphereParams(A, OM, G)
    hs = gaussian_orography(OracleTransforms(o, p), 8, 2)
    for m in (s, o):
        m.set_orography_grid(hs)
        m.set_axis_tilt(0.3)
        m.set_diffusion(1e15)
        m.set_state_grid(h - hs, u, v)
        m.perturb(5.0, 9, 10)
    assert max_rel(s.get_spectral(2), o.get_spectral(2)) <= 1e-12
    s.steps(600.0, 20)
    for _ in range(20):
        o.step(600.0)
    assert rel_l2(s.grid_height(), o.grid_height()) <= 1e-13
    assert rel_l2(s.grid_v(), o.grid_v()) <= 1e-11
    assert abs(s.potential_energy() - o.potential_energy()) <= 1e-12 * o.potential_energy()


def test_run_matches_and_rest_state(lmw, oracle):
    """tests/core.cpp:84-100: a state at rest stays at rest; mass is conserved; run splits like the reference."""
    p = lmw.SphereParams()
    s = lmw.Sphere(10, p)
    o = oracle.OracleSphere(10)
    h = np.full((s.nlat(), s.nlon()), 5000.0)
    z = np.zeros_like(h)
    for m in (s, o):
        m.set_state_grid(h, z, z)
        m.run(86400.0, 1800.0)
    assert s.max_wind() <= 1e-9
    assert abs(s.height_at(0.4, 1.0) - 5000.0) <= 1e-8
    for m in (s, o):
        m.perturb(20.0, 3, 5)
    m0 = s.mean_height()
    s.run(43200.0, 1800.0)
    o.run(43200.0, 1800.0)
    assert abs(s.mean_height() - m0) <= 1e-9
    assert abs(s.time() - o.time()) == 0.0
    assert rel_l2(s.grid_height(), o.grid_height()) <= 1e-13
    with pytest.raises(ValueError, match="max_wave must be between 1 and the truncation"):
        s.perturb(1.0, 1, 0)
    with pytest.raises(ValueError, match="truncation must be between 4 and 170"):
        lmw.Sphere(2)
