"""Mixed-precision Legendre transforms against FP64: one-day ensembles of 16 T85 members (the member-minor state
layout of mixed groups) and one T42 tendency evaluation per mode.  Bounds: FP32 and BF16X6 (FP32-exact three-part
BF16 operands) to 3e-7 relative L2 in h after a day, BF16X3 (two-part operands, ~1e-5 per product) to 1e-5; the
mean height (mass) is conserved exactly in every mode."""
import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

EARTH = (6.371e6, 7.292e-5, 9.80616)
BOUNDS_1DAY = {1: 3e-7, 2: 1e-5, 3: 3e-7}  # FP32, BF16X3, BF16X6
BOUNDS_TEND = {1: 1e-6, 2: 5e-5, 3: 1e-6}


@pytest.fixture(scope="module")
def lmw():
    import lmweather_amd
    return lmweather_amd


def ensemble(lmw, precision, n=16, N=85):
    from lmweather_amd.initial import embed, perturb_vorticity, random_balanced_state
    base = lmw.Sphere(42, lmw.SphereParams(*EARTH))
    h, u, v = random_balanced_state(base, 21, 1)
    base.set_state_grid(h, u, v)
    spec = [base.get_spectral(f) for f in range(3)]
    e = lmw.Ensemble([N] * n, lmw.SphereParams(*EARTH), precision)
    for i in range(n):
        for f in range(3):
            e.set_spectral(i, f, embed(spec[f], N))
        e.set_spectral(i, 0, perturb_vorticity(e.get_spectral(i, 0), 1e-3, 1000 + i))
    return e


@pytest.fixture(scope="module")
def reference(lmw):
    e = ensemble(lmw, lmw.FP64)
    m0 = [e.diag(i, lmw._lib.DIAG_MEAN_HEIGHT) for i in range(16)]
    e.steps([300.0], [288])
    return e, m0


@pytest.mark.parametrize("precision", [1, 2, 3])
def test_one_day_ensemble_against_fp64(lmw, reference, precision):
    ref, m0 = reference
    e = ensemble(lmw, precision)
    # the state went through the member-minor relayout unchanged
    assert np.array_equal(e.get_spectral(5, 2), ensemble(lmw, lmw.FP64).get_spectral(5, 2))
    e.steps([300.0], [288])
    errs = [rel_l2(e.grid(i, 0), ref.grid(i, 0)) for i in range(16)]
    print(f"precision {precision}: rel-L2(h) after 1 day, worst member {max(errs):.2e}")
    assert max(errs) <= BOUNDS_1DAY[precision]
    for i in range(16):
        assert e.diag(i, lmw._lib.DIAG_MEAN_HEIGHT) == m0[i]


@pytest.mark.parametrize("precision", [1, 2, 3])
def test_tendencies_against_fp64(lmw, precision):
    from lmweather_amd.initial import random_balanced_state
    p = lmw.SphereParams(*EARTH)
    a, b = lmw.Sphere(42, p), lmw.Sphere(42, p)
    h, u, v = random_balanced_state(a, 30, 4)
    for s in (a, b):
        s.set_state_grid(h, u, v)
    b.set_precision(precision)
    ta, tb = a.tendencies(), b.tendencies()
    for f, (x, y) in enumerate(zip(tb, ta)):
        e = rel_l2(x, y)
        print(f"precision {precision} field {f}: tendency rel-L2 {e:.2e}")
        # d(delta)/dt and d(phi)/dt of a balanced state are small residuals of nearly cancelling terms: their relative
        # error is that of the terms times the cancellation
        assert e <= BOUNDS_TEND[precision] * (1 if f == 0 else 30)
