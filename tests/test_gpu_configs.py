"""The measured workloads themselves, built exactly as bench.py builds them, under parity tests through the C ABI.

* The headline batch: 32 T170 members (C2 state over the 8 seeded mountains, 1e-3 vorticity perturbations, seeds
  1000 + member) in BF16x6 with the autotuner's tile shapes, i.e. the member-minor layout, the XCD block orders and
  the shared orography plane that the benchmark times.  Eight graph-replayed steps against a 32-member FP64 batch
  (rel-L2 of h <= 1e-8 for every member, mass exact), and member 0 of the FP64 batch against the CPU oracle
  (<= 1e-12).  The other mixed modes at T170 against FP64 with their own bounds.
* C3: one 600 s macro step of the 256-member uneven ensemble (128 x T42, 96 x T85, 32 x T170), in FP64 and BF16x6:
  every member bit for bit equal to its truncation group stepped alone, sampled FP64 members against the oracle and
  sampled BF16x6 members against the FP64 ensemble.
* C4: the 1024 x 512 beta-plane channel with the sech^2 jet and white height noise, a 64-member FP64 batch for three
  steps of stable_dt(0.5): the first and last members against the oracle, FP32 against FP64, mass against the
  oracle's.

Tolerances: FP64 against the oracle as for every other FP64 state (1e-12 relative L2 in h); BF16x6 against FP64 at
1e-8 after 8 steps (its one-day error is ~3e-8); FP32 1e-8 and BF16x3 (16-bit operands) 1e-6.
"""
import numpy as np
import pytest

import bench
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

T170_BOUND_8_STEPS = {1: 1e-8, 2: 1e-6, 3: 1e-8}  # FP32, BF16X3, BF16X6


@pytest.fixture(scope="module")
def lmw():
    import lmweather_amd
    return lmweather_amd


@pytest.fixture(scope="module")
def c2_state(lmw):
    return bench.initial_state(lmw)


@pytest.fixture(scope="module", autouse=True)
def tile_cache(tmp_path_factory):
    """The autotuner's tile cache in a fresh file: the batches are tuned here, as on a fresh machine."""
    import os
    old = os.environ.get("LMW_TUNE_CACHE")
    os.environ["LMW_TUNE_CACHE"] = str(tmp_path_factory.mktemp("tiles") / "tiles.txt")
    yield
    if old is None:
        del os.environ["LMW_TUNE_CACHE"]
    else:
        os.environ["LMW_TUNE_CACHE"] = old


@pytest.fixture(scope="module")
def fp64_batch(lmw, c2_state):
    e = bench.build_batch(lmw, 32, lmw.FP64, c2_state)
    m0 = [e.diag(i, lmw._lib.DIAG_MEAN_HEIGHT) for i in range(32)]
    start = [e.get_spectral(0, f) for f in range(3)]
    e.steps([bench.DT], [8])
    return e, m0, start


def test_fp64_batch_member_matches_oracle(lmw, oracle, c2_state, fp64_batch):
    """Member 0 of the benchmark's FP64 batch after 8 steps against the oracle started from the same spectral state
    (the C2 state plus member 0's perturbation) over the same orography."""
    e, _, start = fp64_batch
    h, u, v, hs = c2_state
    o = oracle.OracleSphere(bench.N, bench.A, bench.OM, bench.G)
    o.set_orography_grid(hs)
    o.set_state_grid(h, u, v)
    for f in range(3):
        o.set_spectral(f, start[f])
    for _ in range(8):
        o.step(bench.DT)
    err = rel_l2(e.grid(0, 0), o.grid_height())
    print(f"FP64 batch member 0 vs oracle after 8 steps: rel-L2(h) {err:.2e}")
    assert err <= 1e-12
    for f in range(3):
        assert rel_l2(e.get_spectral(0, f), o.get_spectral(f)) <= 1e-11, f


@pytest.mark.parametrize("precision", [3, 1, 2])
def test_headline_batch_against_fp64_batch(lmw, c2_state, fp64_batch, precision):
    """The timed batch (autotuned tiles, member-minor layout) stepped 8 times by graph replay: every member against
    the same member of the FP64 batch."""
    ref, m0, _ = fp64_batch
    e = bench.build_batch(lmw, 32, precision, c2_state)
    e.steps([bench.DT], [8])
    errs = np.array([rel_l2(e.grid(i, 0), ref.grid(i, 0)) for i in range(32)])
    eu = max(rel_l2(e.grid(i, 1), ref.grid(i, 1)) for i in (0, 31))
    print(f"precision {precision}: 32 x T170 after 8 steps, rel-L2(h) worst {errs.max():.2e} mean {errs.mean():.2e}; "
          f"u (members 0, 31) {eu:.2e}")
    assert errs.max() <= T170_BOUND_8_STEPS[precision]
    for i in range(32):
        assert e.diag(i, lmw._lib.DIAG_MEAN_HEIGHT) == m0[i]


def c3_members(lmw, precision):
    spec = bench.ensemble_state(lmw)
    return spec, bench.build_uneven(lmw, precision, spec)


@pytest.mark.parametrize("precision", [0, 3])
def test_c3_macro_step_equals_groups_alone(lmw, oracle, precision):
    spec, full = c3_members(lmw, precision)
    groups = bench.ENS_GROUPS
    dts = [dt for _, _, dt in groups]
    subs = [int(round(600.0 / dt)) for dt in dts]
    truncs = [t for t, n, _ in groups for _ in range(n)]
    start = [[full.get_spectral(i, f) for f in range(3)] for i in range(len(truncs))]
    full.macro_steps(dts, subs, 1)
    first = 0
    for (t, n, dt), sub in zip(groups, subs):
        alone = lmw.Ensemble([t] * n, lmw.SphereParams(bench.A, bench.OM, bench.G), precision)
        for i in range(n):
            for f in range(3):
                alone.set_spectral(i, f, start[first + i][f])
        alone.steps([dt], [sub])
        for i in range(n):
            for f in range(3):
                assert np.array_equal(alone.get_spectral(i, f), full.get_spectral(first + i, f)), (t, i, f)
        assert full.diag(first, lmw._lib.DIAG_TIME) == 600.0
        first += n
    if precision == 0:
        for i in (0, 127, 128, 223, 224, 255):
            t = truncs[i]
            dt = dts[[g[0] for g in groups].index(t)]
            o = oracle.OracleSphere(t, bench.A, bench.OM, bench.G)
            for f in range(3):
                o.set_spectral(f, start[i][f])
            for _ in range(int(round(600.0 / dt))):
                o.step(dt)
            err = rel_l2(full.grid(i, 0), o.grid_height())
            print(f"C3 FP64 member {i} (T{t}) vs oracle after 600 s: rel-L2(h) {err:.2e}")
            assert err <= 1e-12, i
    else:
        _, ref = c3_members(lmw, 0)
        ref.macro_steps(dts, subs, 1)
        worst = max(rel_l2(full.grid(i, 0), ref.grid(i, 0)) for i in (0, 127, 128, 223, 224, 255))
        print(f"C3 BF16x6 sampled members vs FP64 after one macro step: rel-L2(h) worst {worst:.2e}")
        assert worst <= 1e-8


def c4_plane(lmw, members, precision):
    nx, ny, lx, ly, g, H, f0, beta = 1024, 512, 1.024e7, 5.12e6, 9.81, 1000.0, 1e-4, 1.6e-11
    from lmweather_amd.initial import normals
    pl = lmw.Plane(lmw.PlaneParams(nx, ny, lx, ly, g, H, f0, beta), members, precision)
    h, u = bench.plane_jet_state(nx, ny, lx, ly, g, H, f0, beta)
    hs = []
    for b in range(members):
        hb = h + 1e-3 * normals(3 + b, nx * ny).reshape(ny, nx)
        hs.append(hb)
        pl.set_field(0, hb, b)
        pl.set_field(1, u, b)
    return pl, hs, u


def test_c4_batch_of_64_against_oracle(lmw, oracle):
    """Three RK4 steps of dt = stable_dt(0.5) of the 64-member FP64 batch: members 0 and 63 against the oracle, the
    mean depth (mass) equal to the oracle's, and the FP32 batch against FP64."""
    pl, hs, u = c4_plane(lmw, 64, lmw.FP64)
    dt = pl.stable_dt(0.5)
    m0 = [pl.mean_depth(b) for b in (0, 63)]
    pl.steps(dt, 3)
    for b, m in zip((0, 63), m0):
        q = oracle.OraclePlane(1024, 512, 1.024e7, 5.12e6, 9.81, 1000.0, 1e-4, 1.6e-11)
        q.set_field(0, hs[b])
        q.set_field(1, u)
        qm0 = q.mean_depth()
        for _ in range(3):
            q.step(dt)
        for f in range(3):
            want = q.get_field(f)
            err = float(np.max(np.abs(pl.get_field(f, b) - want)) / np.max(np.abs(want)))
            assert err <= 1e-13, (b, f, err)
        drift, odrift = pl.mean_depth(b) - m, q.mean_depth() - qm0
        print(f"C4 member {b}: mass drift {drift:.3e} m (oracle {odrift:.3e})")
        assert abs(drift) <= abs(odrift) + 1e-12
    f32, _, _ = c4_plane(lmw, 64, lmw.FP32)
    f32.steps(dt, 3)
    for b in (0, 63):
        dh = rel_l2(f32.get_field(0, b) - 1000.0, pl.get_field(0, b) - 1000.0)
        du = rel_l2(f32.get_field(1, b), pl.get_field(1, b))
        print(f"C4 FP32 member {b} vs FP64: rel-L2(h - H) {dh:.2e}, rel-L2(u) {du:.2e}")
        assert dh <= 1e-3 and du <= 1e-4
