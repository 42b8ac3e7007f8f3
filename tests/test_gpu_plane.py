"""GPU parity of the C-grid plane (reference cpp/plane.cpp) against the CPU oracle, through the C ABI.

FP64 runs with FMA contraction off in the stencil and the RK4 update, so the tendency and the state agree with the
oracle to the last few ulps (the arithmetic and its order are the reference's); FP32 is checked against FP64.
"""
import math

import numpy as np
import pytest

from tests.helpers import max_rel, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lmw():
    import lmweather_amd
    return lmweather_amd


def pair(lmw, oracle, nx=64, ny=32, lx=6.4e5, ly=3.2e5, depth=1000.0, f0=1e-4, beta=1.6e-11, visc=0.0, drag=0.0, linear=False,
         batch=1, precision=0):
    p = lmw.PlaneParams(nx, ny, lx, ly, 9.81, depth, f0, beta, visc, drag, linear)
    return lmw.Plane(p, batch, precision), oracle.OraclePlane(nx, ny, lx, ly, 9.81, depth, f0, beta, visc, drag, linear)


def seed_state(g, o, seed=0, member=0):
    rng = np.random.default_rng(seed)
    nx, ny = o.nx, o.ny
    h = 1000.0 + rng.standard_normal((ny, nx))
    u = 0.5 * rng.standard_normal((ny, nx))
    v = 0.5 * rng.standard_normal((ny + 1, nx))
    for m, kw in ((g, {"member": member}), (o, {})):
        m.set_field(0, h, **kw)
        m.set_field(1, u, **kw)
        m.set_field(2, v, **kw)
    return h, u, v


@pytest.mark.parametrize("visc,drag,linear", [(0.0, 0.0, False), (50.0, 1e-6, False), (10.0, 0.0, True)])
def test_tendency(lmw, oracle, visc, drag, linear):
    g, o = pair(lmw, oracle, visc=visc, drag=drag, linear=linear)
    seed_state(g, o, 1)
    for a, b in zip(g.tendency(), o.tendency()):
        assert max_rel(a, b) <= 1e-14


def test_steps_fp64(lmw, oracle):
    g, o = pair(lmw, oracle, visc=20.0, drag=1e-7)
    seed_state(g, o, 2)
    g.steps(30.0, 100)
    for _ in range(100):
        o.step(30.0)
    for f in range(3):
        assert max_rel(g.get_field(f), o.get_field(f)) <= 1e-13
    assert abs(g.energy() - o.energy()) <= 1e-12 * o.energy()
    assert abs(g.mean_depth() - o.mean_depth()) <= 1e-13 * o.mean_depth()
    assert abs(g.max_speed() - o.max_speed()) <= 1e-12 * o.max_speed()
    assert abs(g.potential_vorticity_anomaly_rms() - o.pv_anomaly_rms()) <= 1e-11 * o.pv_anomaly_rms()
    assert abs(g.stable_dt() - o.stable_dt()) <= 1e-12 * o.stable_dt()


def test_batch_members_independent(lmw, oracle):
    g, o = pair(lmw, oracle, batch=3)
    seed_state(g, o, 5, member=2)
    g.steps(30.0, 10)
    for _ in range(10):
        o.step(30.0)
    assert max_rel(g.get_field(0, member=2), o.get_field(0)) <= 1e-13
    assert np.all(g.get_field(0, member=0) == 1000.0)  # the other members started at rest (reset) and stay there


def test_fp32_against_fp64(lmw, oracle):
    g64, o = pair(lmw, oracle)
    g32, _ = pair(lmw, oracle, precision=1)
    h, u, v = seed_state(g64, o, 3)
    for f, x in enumerate((h, u, v)):
        g32.set_field(f, x)
    g64.steps(30.0, 50)
    g32.steps(30.0, 50)
    assert rel_l2(g32.get_field(0) - 1000.0, g64.get_field(0) - 1000.0) <= 1e-3
    assert rel_l2(g32.get_field(1), g64.get_field(1)) <= 1e-4


def test_gravity_wave_and_errors(lmw, oracle):
    """tests/core.cpp:101-117: a linear standing gravity wave after half a period; mass; bad parameters."""
    p = lmw.PlaneParams(100, 8, 1.0e6, 2.0e4, 9.81, 100.0, 0.0, 0.0, 0.0, 0.0, True)
    pl = lmw.Plane(p)
    k, c = 2 * math.pi / p.lx, math.sqrt(p.gravity * p.depth)
    h = np.array([[p.depth + math.cos(k * pl.x_center(i)) for i in range(p.nx)] for _ in range(p.ny)])
    pl.set_h_from_grid(h)
    pl.run(math.pi / (c * k), 0)
    assert abs(pl.h_at(0.5 * pl.dx(), 1e4) - p.depth + math.cos(k * 0.5 * pl.dx())) <= 2e-3
    assert abs(pl.mean_depth() - p.depth) <= 1e-10
    with pytest.raises(ValueError, match="grid must be at least 4 x 4"):
        lmw.Plane(lmw.PlaneParams(2, 2, 1.0, 1.0))
    g, o = pair(lmw, oracle)
    seed_state(g, o, 4)
    with pytest.raises(ValueError) as e1:
        g.run(1e5, 1e4)
    with pytest.raises(ValueError) as e2:
        o.run(1e5, 1e4)
    assert str(e1.value) == str(e2.value)


def test_run_matches_oracle(lmw, oracle):
    g, o = pair(lmw, oracle, visc=5.0)
    seed_state(g, o, 6)
    o.perturb(0.5, 9)
    g.perturb(0.5, 9)
    assert max_rel(g.get_field(0), o.get_field(0)) == 0.0
    g.run(3600.0 * 3, 0)
    o.run(3600.0 * 3, 0)
    assert g.time() == o.time()
    assert max_rel(g.get_field(0), o.get_field(0)) <= 1e-13


@pytest.mark.parametrize("precision,tiles,physics", [(0, (5, 0, 1, 2, 4), "viscous"), (1, (5, 0, 1, 3), "viscous"),
                                                     (0, (5, 0, 1), "linear"), (0, (5, 0, 1, 4), "inviscid"),
                                                     (1, (5, 0, 1, 2), "inviscid")])
def test_fused_step_tiles_bit_identical(lmw, oracle, precision, tiles, physics):
    """The row march (one launch per RK4 step) in chunks of 32 to 256 rows against four stage launches: bit-identical
    results; FP64 equal to the oracle to the last bits on a grid whose width is not a multiple of the 56-column strips
    (a partial strip that wraps around) and whose 37 rows split into 32-row chunks (the second starts 4 rows early)."""
    nx, ny = 100, 37
    visc, drag, linear = {"viscous": (50.0, 1e-6, False), "linear": (0.0, 0.0, True), "inviscid": (0.0, 1e-6, False)}[physics]
    p = lmw.PlaneParams(nx, ny, 1.0e6, 3.7e5, 9.81, 1000.0, 1e-4, 1.6e-11, visc, drag, linear)
    rng = np.random.default_rng(5)
    h = 1000.0 + rng.standard_normal((ny, nx))
    u = 0.5 * rng.standard_normal((ny, nx))
    v = 0.5 * rng.standard_normal((ny + 1, nx))
    v[0] = v[-1] = 0.0
    outs = []
    for t in tiles:
        g = lmw.Plane(p, 2, precision)
        g.set_tile(t)
        for m in range(2):
            g.set_field(0, h + m, m)
            g.set_field(1, u, m)
            g.set_field(2, v, m)
        g.steps(20.0, 7)
        outs.append([g.get_field(f, 1) for f in range(3)])
    for t, o in zip(tiles[1:], outs[1:]):
        for f, (a, b) in enumerate(zip(o, outs[0])):
            assert np.array_equal(a, b), (t, f, np.max(np.abs(a - b)))
    if precision == 0:
        q = oracle.OraclePlane(nx, ny, 1.0e6, 3.7e5, 9.81, 1000.0, 1e-4, 1.6e-11, visc, drag, linear)
        q.set_field(0, h + 1)
        q.set_field(1, u)
        q.set_field(2, v)
        for _ in range(7):
            q.step(20.0)
        for f in range(3):
            ref = q.get_field(f)
            assert np.max(np.abs(outs[0][f] - ref)) <= 1e-12 * np.max(np.abs(ref)), f


@pytest.mark.parametrize("precision,nx,ny", [(0, 64, 32), (0, 1024, 512), (1, 200, 48)])
def test_perturb_bit_identical(lmw, oracle, precision, nx, ny):
    """perturb (plane.cpp:34-57) generates the seeded noise and smooths it on the device; the result is the
    reference's bit for bit (FP64), and its rounding to FP32 for an FP32 plane.  Members are independent."""
    g, o = pair(lmw, oracle, nx=nx, ny=ny, lx=nx * 1e4, ly=ny * 1e4, batch=3, precision=precision)
    g.perturb(0.75, 12345, member=1)
    o.perturb(0.75, 12345)
    want = o.get_field(0)
    got = g.get_field(0, member=1)
    if precision == 0:
        assert np.array_equal(got, want)
    else:
        assert np.array_equal(got, want.astype(np.float32).astype(np.float64))
    assert np.all(g.get_field(0, member=0) == 1000.0) and np.all(g.get_field(0, member=2) == 1000.0)
    g.perturb(0.75, 12345, member=1)  # a second perturbation adds another field of the same stream
    o.perturb(0.75, 12345)
    if precision == 0:
        assert np.array_equal(g.get_field(0, member=1), o.get_field(0))


@pytest.mark.parametrize("nx,ny", [(4, 4), (5, 7), (57, 33), (100, 129), (200, 257), (1000, 6), (4000, 20)])
def test_odd_sizes_match_oracle(lmw, oracle, nx, ny):
    """Grids that are not multiples of the march's 56-column strips or 128-row chunks, the smallest grid the
    reference accepts (4 x 4) and the widest row (4000): nonlinear, viscous, with drag, FP64 against the oracle and
    FP32 against FP64."""
    lx, ly = nx * 1e4, ny * 1e4
    g, o = pair(lmw, oracle, nx=nx, ny=ny, lx=lx, ly=ly, visc=30.0, drag=1e-6)
    f, _ = pair(lmw, oracle, nx=nx, ny=ny, lx=lx, ly=ly, visc=30.0, drag=1e-6, precision=1)
    seed_state(g, o, nx + ny)
    for field in range(3):
        f.set_field(field, g.get_field(field))
    dt = 0.5 * o.stable_dt()
    g.steps(dt, 5)
    f.steps(dt, 5)
    for _ in range(5):
        o.step(dt)
    for field in range(3):
        want = o.get_field(field)
        assert max_rel(g.get_field(field), want) <= 1e-13, field
        assert max_rel(f.get_field(field), g.get_field(field)) <= 1e-5, field
