"""GPU error paths of the reference's script tests (test.sh:176-215) that the model library owns, set up through the
same operators the scripts use: the texts of sphere.cpp / plane.cpp (the depth check, the run refusals, the blow-up
and the mid-run stability recheck, the crest and domain checks), against the oracle driven through the same
sequence where the text carries numbers.
"""
import math

import pytest

from oracle import expr as X

pytestmark = pytest.mark.gpu

A, OM, G = 6.37122e6, 7.292e-5, 9.80616


@pytest.fixture(scope="module")
def lmw():
    import lmweather_amd
    return lmweather_amd


def sphere(lmw, N):
    return lmw.Sphere(N, lmw.SphereParams(A, OM, G))


def test_zero_height_and_crest(lmw):
    """test.sh:184 zero_height, :201 bad_crest."""
    s = sphere(lmw, 10)
    s.fields("0", "0", "0")
    with pytest.raises(RuntimeError, match=r"^the height field must be positive to step the model$"):
        s.stable_dt()
    t = sphere(lmw, 10)
    t.fields("1000", "0", "0")
    with pytest.raises(ValueError, match=r"^zonal wavenumber must be between 1 and the truncation$"):
        t.wave_crest(0, math.radians(45.0))


def test_plane_refusals(lmw, oracle):
    """test.sh:178 bad_domain, :185 unstable_dt, :187 blow_up, :202 no_rossby_radius."""
    with pytest.raises(ValueError, match=r"^domain sizes must be positive$"):
        lmw.Plane(lmw.PlaneParams(10, 10, 0.0, 1.0))
    with pytest.raises(ValueError, match=r"^Plane.rossby_radius: f0 is zero$"):
        lmw.Plane(lmw.PlaneParams(10, 10, 1e6, 5e5)).rossby_radius()
    for h, seconds, dt, kind in [("cos(x / 100000.0)", 5e6, 5e4, ValueError), ("-150.0 * cos(2.0 * pi * x / lx)", 5e7, 0.0, RuntimeError)]:
        p = lmw.Plane(lmw.PlaneParams(20, 10, 1e6, 5e5, depth=100.0))
        p.fields(h, "0", "0")
        o = oracle.OraclePlane(20, 10, 1e6, 5e5, depth=100.0)
        for f, g in enumerate(X.plane_fields(h, "0", "0", 20, 10, 1e6, 5e5, 100.0)):
            o.set_field(f, g)
        with pytest.raises(kind) as e1:
            p.run(seconds, dt)
        with pytest.raises(kind) as e2:
            o.run(seconds, dt)
        if kind is ValueError:  # the refusal before the first step: the same text, numbers included
            assert str(e1.value) == str(e2.value) and "exceeds the RK4 stability limit" in str(e1.value)
        else:
            assert str(e1.value) == str(e2.value) == "the integration blew up (not finite): use a smaller dt"


def test_sphere_refusals(lmw, oracle):
    """test.sh:186 unstable_sphere_dt and :210 accelerated: refused before the first step, and the recheck every 64
    steps that catches a flow whose winds spun up (the same step and time as the oracle)."""
    s = sphere(lmw, 21)
    s.setup(2, 0.0)
    with pytest.raises(ValueError, match="exceeds the RK4 stability limit"):
        s.run(86400.0, 86400.0)
    h = "5960.0 - 1500.0 * pow(2.718281828459045, -((lon - 3.0) * (lon - 3.0) * cos(lat) * cos(lat) + lat * lat) / 0.1)"
    s = sphere(lmw, 21)
    s.fields(h, "0", "0")
    o = oracle.OracleSphere(21, A, OM, G)
    mu = o.gauss()[0]
    o.set_state_grid(*X.sphere_fields(h, "0", "0", mu, o.nlon, X.sphere_params(A, OM, G)))
    dt = 0.97 * s.stable_dt() * 1.8856180831641
    assert dt == pytest.approx(0.97 * o.stable_dt() * 1.8856180831641, rel=1e-12)
    with pytest.raises(RuntimeError) as e1:
        s.run(5.0 * 86400.0, dt)
    with pytest.raises(RuntimeError) as e2:
        o.run(5.0 * 86400.0, dt)
    for e in (e1, e2):
        assert "(the flow accelerated): use a smaller dt" in str(e.value)
    # the same substep length and the same time reached; the limit's last printed digit may differ by rounding
    assert str(e1.value).split(" exceeds")[0] == str(e2.value).split(" exceeds")[0]
    assert str(e1.value).split("reached at")[1] == str(e2.value).split("reached at")[1]
