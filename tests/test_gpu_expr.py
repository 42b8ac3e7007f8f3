"""GPU expression grids, through the C ABI: the Sphere.fields / orography / tracer and Plane.fields operators
(cpp/main.cpp:158-172,212-228,252-268,410-431) compiled on the host and evaluated on the device, against the oracle's
restatement of the reference's compiler (oracle/expr.py) feeding the oracle model, plus the reference's own script
tests that use them (tests/sphere_spectrum.lms, sphere_orography.lms, plane_rossby.lms, plane_adjustment.lms,
test.sh:179-215).
"""
import numpy as np
import pytest

from oracle import expr as X
from tests.helpers import max_rel

pytestmark = pytest.mark.gpu

A, OM, G = 6.37122e6, 7.292e-5, 9.80616
PAR = X.sphere_params(A, OM, G)
CONE = "2000.0 * (1 - sqrt(min((pi / 9) * (pi / 9), (lon - 1.5 * pi) * (lon - 1.5 * pi) + (lat - pi / 6) * (lat - pi / 6))) / (pi / 9))"
FIELDS = [
    (21, "2000.0", "15.0 * cos(lat)", "6.0 * cos(3.0 * lon) * cos(lat) * cos(lat)"),  # sphere_spectrum.lms:8
    (42, "5960.0 - 100.0 * pow(2.718281828459045, -((lon * cos(lat)) * (lon * cos(lat)) + (lat - 0.7853981633974483) * "
         "(lat - 0.7853981633974483)) / 0.04)", "0", "0"),  # sphere_adjustment.lms:18
    (21, "29400.0 / g - (a * omega * 38.6106 + 38.6106 * 38.6106 / 2.0) * pow(0 - cos(lon) * cos(lat) * sin(1.5) + "
         "sin(lat) * cos(1.5), 2) / g", "38.6106 * (cos(lat) * cos(1.5) + sin(lat) * cos(lon) * sin(1.5))",
     "0 - 38.6106 * sin(lon) * sin(1.5)"),  # sphere_williamson2.lms:32
    (85, "5960.0 - " + CONE, "20 * cos(lat) + tanh(lon - pi) + sign(lat) * abs(lat)^1.5", "mod(lon, 1.0) * max(0, lat) - atan2(lat, lon + 1)"),
]


@pytest.fixture(scope="module")
def lmw():
    import lmweather_amd
    return lmweather_amd


def sphere(lmw, N, precision=0):
    return lmw.Sphere(N, lmw.SphereParams(A, OM, G), precision)


def oracle_grids(o, *texts):
    mu = o.gauss()[0]
    return [X.sphere_grid(t, mu, o.nlon, PAR) for t in texts]


@pytest.mark.parametrize("case", range(len(FIELDS)))
def test_sphere_fields_match_oracle(lmw, oracle, case):
    N, h, u, v = FIELDS[case]
    s = sphere(lmw, N)
    s.fields(h, u, v)
    o = oracle.OracleSphere(N, A, OM, G)
    o.set_state_grid(*oracle_grids(o, h, u, v))
    for f in range(3):
        sc = np.max(np.abs(o.get_spectral(0 if f == 1 else f)))
        assert max_rel(s.get_spectral(f), o.get_spectral(f), scale=sc) <= 1e-12, f
    assert s.height_error_l2() == 0.0  # set_state_grid marks the reference


def test_orography_and_tracer_match_oracle(lmw, oracle):
    s = sphere(lmw, 42)
    s.orography(CONE)
    s.fields("5960.0 - " + CONE, "0", "0")
    s.tracer("500.0 * (1 + cos(pi * min(1.0, 3.0 * acos(cos(lat) * cos(lon - 1.5 * pi)))))")  # sphere_tracer.lms:38
    o = oracle.OracleSphere(42, A, OM, G)
    hs, h, q = oracle_grids(o, CONE, "5960.0 - " + CONE, "500.0 * (1 + cos(pi * min(1.0, 3.0 * acos(cos(lat) * cos(lon - 1.5 * pi)))))")
    assert max_rel(s.grid_orography(), hs, scale=2000.0) <= 1e-14
    o.set_orography_grid(hs)
    o.set_state_grid(h, np.zeros_like(h), np.zeros_like(h))
    o.set_tracer_grid(q)
    assert max_rel(s.get_spectral(3), o.get_spectral(3)) <= 1e-12
    for f in (0, 2):
        assert max_rel(s.get_spectral(f), o.get_spectral(f), scale=np.max(np.abs(o.get_spectral(f)))) <= 1e-12
    sc = [2 * OM * np.max(np.abs(o.get_spectral(2)))] * 3
    for a, b, c in zip(s.tendencies(), o.tendencies(), sc):
        assert np.max(np.abs(a - b)) <= 1e-11 * c


def test_script_known_answers(lmw):
    """sphere_spectrum.lms:8-16 and sphere_orography.lms:8-21 through the expression operators."""
    s = sphere(lmw, 21)
    s.fields("2000.0", "15.0 * cos(lat)", "6.0 * cos(3.0 * lon) * cos(lat) * cos(lat)")
    assert abs(s.energy() - (s.kinetic_energy() + s.potential_energy())) <= 1e-6 * s.energy()
    assert abs(s.potential_energy() - 0.5 * G * 2000.0 * 2000.0) <= 1e-6
    r = sphere(lmw, 21)
    r.fields("2000.0", "15.0 * cos(lat)", "0")
    assert abs(r.kinetic_energy() / 2000.0 - 15.0 * 15.0 / 3.0) <= 1e-7
    assert r.kinetic_energy() < s.kinetic_energy()
    rest = sphere(lmw, 21)
    rest.orography(CONE)
    rest.fields("5960.0 - " + CONE, "0", "0")
    e0 = rest.energy()
    assert abs(rest.height_at(np.radians(30.0), np.radians(270.0)) - 3960.0) <= 300.0
    assert abs(rest.height_at(np.radians(30.0), np.radians(90.0)) - 5960.0) <= 1.0
    rest.run(86400.0, 1200.0)
    assert rest.max_wind() < 1e-9 and rest.height_error_l2() < 1e-12 and abs(rest.energy() / e0 - 1.0) <= 1e-12
    hole = sphere(lmw, 21)
    hole.fields("5960.0 - " + CONE, "0", "0")
    hole.run(86400.0, 1200.0)
    assert hole.max_wind() > 10.0 and hole.height_error_l2() > 0.01


BAD = ["2 +* 3", "foo", "foo(1)", "sin(1, 2)", "(1 + 2", "", "1e400", "lat)", "(" * 300 + "1" + ")" * 300, "x + 1",
       "+".join(["1"] * 2001)]


def _oracle_error(text, vars, params):
    try:
        X.Expr(text, vars, params)
    except X.ExprError as e:
        return str(e)
    return None


def test_expression_errors(lmw):
    """Compile errors carry the reference compiler's texts (pinned against it in test_oracle_expr.py); non-finite
    grids the operators' texts (main.cpp:170,226,265,425-427, test.sh:212-214); a failed call leaves the state as
    it was."""
    s = sphere(lmw, 21)
    s.fields("5960", "10 * cos(lat)", "0")
    before = [s.get_spectral(f) for f in range(3)]
    hs0 = s.grid_orography()
    for t in BAD:
        want = _oracle_error(t, ["lon", "lat"], PAR)
        assert want is not None, t
        for call in (lambda: s.fields("1000", t, "0"), lambda: s.orography(t), lambda: s.tracer(t)):
            with pytest.raises(RuntimeError) as e:
                call()
            assert str(e.value) == want, t
    for call, msg in [(lambda: s.fields("1000", "0", "1 / (lat - lat)"), "Sphere.fields: an expression is not finite on the grid"),
                      (lambda: s.fields("1e308", "1e308", "0"), "Sphere.fields: an expression is not finite on the grid"),
                      (lambda: s.orography("1 / (lat - lat)"), "Sphere.orography: the expression is not finite on the grid"),
                      (lambda: s.tracer("1 / (lat - lat)"), "Sphere.tracer: the expression is not finite on the grid")]:
        with pytest.raises(RuntimeError) as e:
            call()
        assert str(e.value) == msg
    for f in range(3):
        assert np.array_equal(s.get_spectral(f), before[f])
    assert np.array_equal(s.grid_orography(), hs0) and not s.has_tracer()
    p = lmw.Plane(lmw.PlaneParams(16, 8, 1e6, 5e5, depth=100.0))
    h0 = p.get_field(0)
    for t in BAD[:-2]:
        want = _oracle_error(t, ["x", "y"], X.plane_params(1e6, 5e5, 100.0, 9.81, 0.0, 0.0))
        with pytest.raises(RuntimeError) as e:
            p.fields("0", "0", t)
        assert str(e.value) == want, t
    for args, what in [(("1 / (x - x)", "0", "0"), "height"), (("0", "log(0 * y)", "0"), "u"), (("0", "0", "1 / (y - y)"), "v"),
                       (("1e308 * 10", "0", "0"), "height")]:
        with pytest.raises(RuntimeError) as e:
            p.fields(*args)
        assert str(e.value) == f"Plane.fields: the {what} expression is not finite"
    assert np.array_equal(p.get_field(0), h0)
    # v is evaluated at the interior faces only: 1 / y is finite there and on neither wall
    p.fields("0", "0", "1 / y")


def test_deep_expressions(lmw, oracle):
    """Expressions whose evaluation stack outgrows the register-resident variant (16 values) run the deep variant;
    99 levels of '1 + (' is the deepest a string may nest (cpp/expr.hpp:51)."""
    o = oracle.OracleSphere(10, A, OM, G)
    for levels, inner in ((20, "lat * cos(lon)"), (99, "lat")):
        t = "1 + (" * levels + inner + ")" * levels
        s = sphere(lmw, 10)
        s.orography(t)
        want = oracle_grids(o, t)[0]
        assert max_rel(s.grid_orography(), want, scale=np.max(np.abs(want))) <= 1e-14, levels
    with pytest.raises(RuntimeError, match="nested too deeply"):
        s.orography("1 + (" * 100 + "lat" + ")" * 100)


@pytest.mark.parametrize("precision", [0, 1])
def test_plane_fields_match_oracle(lmw, precision):
    """plane_rossby.lms:17 and plane_harmonic.lms:11 states; FP32 planes hold the FP64 values rounded once."""
    nx, ny, lx, ly, depth, f0, beta = 60, 30, 6.0e6, 3.0e6, 1000.0, 1e-4, 1.6e-11
    cases = [("f0 * 100000.0 * cos(2 * pi * x / lx) * sin(pi * y / ly) / g", "-100000.0 * pi / ly * cos(2 * pi * x / lx) * cos(pi * y / ly)",
              "-100000.0 * 2 * pi / lx * sin(2 * pi * x / lx) * sin(pi * y / ly)"),
             ("0", "3.0 * cos(2 * pi * x / lx)", "4.0 * sin(4 * pi * x / lx) * sin(pi * y / ly)"),
             ("H * 0.001 * beta * y / f0 + exp(-(y - ly / 2) * (y - ly / 2) / 160000000000.0)", "x / lx", "y / ly")]
    p = lmw.Plane(lmw.PlaneParams(nx, ny, lx, ly, depth=depth, f0=f0, beta=beta), batch=2, precision=precision)
    for h, u, v in cases:
        p.fields(h, u, v, member=1)
        want = X.plane_fields(h, u, v, nx, ny, lx, ly, depth, 9.81, f0, beta)
        for f in range(3):
            got = p.get_field(f, member=1)
            if precision == 0:
                assert max_rel(got, want[f], scale=max(1e-300, np.max(np.abs(want[f])))) <= 1e-14, (h, f)
            else:
                assert max_rel(got, want[f].astype(np.float32).astype(np.float64), scale=max(1e-300, np.max(np.abs(want[f])))) <= 1e-7, (h, f)
    assert np.all(p.get_field(0, member=0) == depth)  # member 0 untouched


def test_ensemble_fields_and_orography(lmw):
    """Members of two truncations set in batched launches equal single spheres given the same expressions; members
    given one orography expression share one plane."""
    truncs = [21, 21, 42, 21, 42]
    e = lmw.Ensemble(truncs, lmw.SphereParams(A, OM, G))
    e.orography(0, 5, CONE)
    e.fields(0, 5, "5960.0 - " + CONE, "10 * cos(lat)", "0")
    for i, N in enumerate(truncs):
        s = sphere(lmw, N)
        s.orography(CONE)
        s.fields("5960.0 - " + CONE, "10 * cos(lat)", "0")
        assert np.array_equal(e.grid(i, 5), s.grid_orography()), i
        for f in range(3):
            assert max_rel(e.get_spectral(i, f), s.get_spectral(f), scale=np.max(np.abs(s.get_spectral(0 if f == 1 else f)))) <= 1e-14, (i, f)
        assert e.diag(i, 3) == pytest.approx(s.potential_energy(), rel=1e-14)


ROSSBY = ("f0 * 100000.0 * cos(2 * pi * x / lx) * sin(pi * y / ly) / g", "-100000.0 * pi / ly * cos(2 * pi * x / lx) * cos(pi * y / ly)",
          "-100000.0 * 2 * pi / lx * sin(2 * pi * x / lx) * sin(pi * y / ly)")


def test_plane_script_known_answers(lmw):
    """plane_rossby.lms (the beta drift of a Rossby wave, none on an f-plane, twice as fast at 2 beta) and the first
    part of plane_adjustment.lms (deformation radius, PV and mass through five inertial periods), on the device."""
    lx, ly, f0, beta = 4.0e6, 2.0e6, 1e-4, 2e-11

    def plane(b):
        p = lmw.Plane(lmw.PlaneParams(60, 30, lx, ly, depth=1000.0, f0=f0, beta=b, linear=True))
        p.fields(*ROSSBY)
        return p

    k, l = 2 * np.pi / lx, np.pi / ly
    m = plane(beta)
    ld = m.rossby_radius()
    cp = -beta / (k * k + l * l + 1.0 / (ld * ld))
    assert abs(cp - -3.359) <= 0.002
    c0, amp0 = m.zonal_crest(1, 15), m.zonal_amplitude(1, 15)
    t = 20.0 * 86400.0
    m.run(t, 0.0)
    diff = m.zonal_crest(1, 15) - c0 - cp * t
    assert abs(diff - np.floor(diff / lx + 0.5) * lx) <= 60000.0
    assert abs(m.zonal_amplitude(1, 15) / amp0 - 1.0) <= 0.1
    f = plane(0.0)
    fc0 = f.zonal_crest(1, 15)
    f.run(t, 0.0)
    moved = f.zonal_crest(1, 15) - fc0
    assert abs(moved - np.floor(moved / lx + 0.5) * lx) < 0.02 * lx
    shifts = []
    for b in (2 * beta, beta):
        p = plane(b)
        c = p.zonal_crest(1, 15)
        p.run(2.0 * 86400.0, 0.0)
        s = p.zonal_crest(1, 15) - c
        shifts.append(s - np.floor(s / lx + 0.5) * lx)
    assert abs(shifts[0] / shifts[1] - 2.0) <= 0.15
    a = lmw.Plane(lmw.PlaneParams(8, 200, 40000.0, 4.0e6, depth=1000.0, f0=f0, beta=0.0, linear=True))
    a.fields("exp(-(y - ly / 2) * (y - ly / 2) / 160000000000.0)", "0", "0")
    assert abs(a.rossby_radius() - 990515.0) <= 100.0
    q0, mean0 = a.potential_vorticity_anomaly_rms(), a.mean_depth()
    a.run(5.0 * 2.0 * np.pi / f0, 0.0)
    assert abs(a.potential_vorticity_anomaly_rms() / q0 - 1.0) <= 0.001
    assert abs(a.mean_depth() - mean0) <= 1e-11
