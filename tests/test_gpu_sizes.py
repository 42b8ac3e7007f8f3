"""GPU parity across the truncation range the reference accepts (4..170, cpp/sphere.cpp:76): the smallest grids (FFT
lengths 16 and 32), odd truncations (odd row counts per parity, padding of the MFMA row tiles) and the largest, in
FP64 against the oracle and in the mixed modes against the GPU's FP64, one sphere at a time and as one uneven
ensemble stepped through its HIP graph.
"""
import numpy as np
import pytest

from tests.helpers import OracleTransforms, rel_l2

pytestmark = pytest.mark.gpu

A, OM, G = 6.371e6, 7.292e-5, 9.80616
SIZES = [4, 5, 7, 10, 15, 31, 63, 106, 127, 169]
DT = 150.0


@pytest.fixture(scope="module")
def lmw():
    import lmweather_amd
    return lmweather_amd


@pytest.fixture(scope="module")
def states(oracle):
    """A balanced random flow per truncation (energy up to n = min(N, 12)), built with the oracle's transforms."""
    from lmweather_amd._lib import SphereParams
    from lmweather_amd.initial import random_balanced_state
    out = {}
    for N in SIZES:
        o = oracle.OracleSphere(N, A, OM, G)
        out[N] = random_balanced_state(OracleTransforms(o, SphereParams(A, OM, G)), min(N, 12), 7 + N)
    return out


@pytest.mark.parametrize("N", SIZES)
def test_fp64_steps_match_oracle(lmw, oracle, states, N):
    h, u, v = states[N]
    s = lmw.Sphere(N, lmw.SphereParams(A, OM, G))
    o = oracle.OracleSphere(N, A, OM, G)
    s.set_state_grid(h, u, v)
    o.set_state_grid(h, u, v)
    s.steps(DT, 3)
    for _ in range(3):
        o.step(DT)
    assert rel_l2(s.grid_height(), o.grid_height()) <= 1e-12
    zs = np.max(np.abs(o.get_spectral(0)))
    for f in range(3):
        assert np.max(np.abs(s.get_spectral(f) - o.get_spectral(f))) <= 1e-12 * (zs if f < 2 else np.max(np.abs(o.get_spectral(2))))


@pytest.mark.parametrize("precision", [1, 2, 3])
def test_mixed_modes_every_size(lmw, states, precision):
    """FP32, BF16x3 and BF16x6 Legendre contractions against the GPU's FP64 after 3 steps (rel-L2 of h)."""
    bound = {1: 1e-8, 2: 1e-6, 3: 1e-8}[precision]
    for N in SIZES:
        h, u, v = states[N]
        a = lmw.Sphere(N, lmw.SphereParams(A, OM, G))
        b = lmw.Sphere(N, lmw.SphereParams(A, OM, G), precision)
        for s in (a, b):
            s.set_state_grid(h, u, v)
            s.steps(DT, 3)
        hb = b.grid_height()
        assert np.all(np.isfinite(hb)), N
        assert rel_l2(hb, a.grid_height()) <= bound, N


@pytest.mark.parametrize("precision", [0, 3])
def test_uneven_ensemble_every_size(lmw, states, precision):
    """One ensemble holding every truncation (ten groups in ascending order): its graph-stepped members equal the
    same states stepped as single spheres."""
    e = lmw.Ensemble(SIZES[::-1], lmw.SphereParams(A, OM, G), precision)
    for i, N in enumerate(SIZES[::-1]):
        e.set_state_grid(i, *states[N])
    e.steps([DT] * len(SIZES), [3] * len(SIZES))
    for i, N in enumerate(SIZES[::-1]):
        s = lmw.Sphere(N, lmw.SphereParams(A, OM, G), precision)
        s.set_state_grid(*states[N])
        s.steps(DT, 3)
        zs = np.max(np.abs(s.get_spectral(0)))
        for f in range(3):
            d = np.max(np.abs(e.get_spectral(i, f) - s.get_spectral(f)))
            assert d <= (1e-13 if precision == 0 else 1e-7) * (zs if f < 2 else np.max(np.abs(s.get_spectral(2)))), (N, f)
