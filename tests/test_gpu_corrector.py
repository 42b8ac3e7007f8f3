"""GPU tests of the learned coarse-step corrector: the hand-written forward, backward and Adam kernels against the
CPU restatement of the network, the device truncation of ensemble states, the corrected step's bookkeeping
(correction, fallback, per-member merge), and what the trained network gains over the coarse step without it."""
import os

import numpy as np
import pytest

from oracle import corrector_net as cn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lmw():
    import lmweather_amd
    return lmweather_amd


@pytest.fixture(scope="module")
def corr():
    from lmweather_amd import corrector
    return corrector


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def _data(N, S, seed):
    rng = np.random.default_rng(seed)
    C = cn.tri_size(N)
    deg = np.concatenate([np.arange(m, N + 1) for m in range(N + 1)])
    amp = (1.0 + deg) ** -1.0  # a red spectrum, like the model's
    x = (rng.standard_normal((S, 3, C, 2)) * amp[None, None, :, None] * 3.0).astype(np.float32)
    t = (rng.standard_normal((S, 3, C, 2)) * amp[None, None, :, None] * 0.05).astype(np.float32)
    ins = np.tile((1.0 + np.arange(N + 1)) ** -1.0 * 3.0, 3).astype(np.float32)
    outs = np.tile((1.0 + np.arange(N + 1)) ** -1.0 * 0.05, 3).astype(np.float32)
    return x, t, ins, outs


@pytest.mark.parametrize("N,W,S", [(10, 16, 3), (42, 64, 2)])
def test_loss_and_gradient_match_oracle(corr, N, W, S):
    net = corr.Corrector(N, W, seed=5)
    x, t, ins, outs = _data(N, S, 11)
    net.set_scales(ins, outs)
    p = net.params
    loss, g = net.loss_grad(x, t)
    lo, go = cn.loss_grad(p, W, N, ins, outs, x, t, round_weights=True)
    print(f"N={N} W={W}: loss {loss:.6e} oracle {lo:.6e}")
    assert abs(loss - lo) <= 2e-5 * lo
    lay, _ = cn.layout(W)
    for name, (o, shp) in lay.items():
        n = int(np.prod(shp))
        e = _rel(g[o:o + n], go[o:o + n])
        print(f"  grad {name}: rel-L2 {e:.2e}")
        assert e <= 2e-4, name


def test_forward_matches_oracle(corr):
    N, W, S = 21, 32, 4
    net = corr.Corrector(N, W, seed=7)
    x, _, ins, outs = _data(N, S, 12)
    net.set_scales(ins, outs)
    c, r = net.forward(x)
    co, ro = cn.correction(net.params, W, N, ins, outs, x)
    assert _rel(c, co) <= 1e-5
    assert np.allclose(r, ro, rtol=1e-5, atol=0)
    assert np.all(c[:, :, 0, :] == 0)


def test_adam_training_matches_oracle(corr):
    N, W, S, lr, steps = 8, 16, 4, 1e-3, 3
    net = corr.Corrector(N, W, seed=9)
    x, t, ins, outs = _data(N, S, 13)
    net.set_scales(ins, outs)
    p0 = net.params.astype(np.float64)
    net.set_dataset(x, t)
    losses = net.train(steps, S, lr, seed=1)  # one full batch per epoch: one Adam step each
    p = p0.copy()
    m1, m2 = np.zeros_like(p), np.zeros_like(p)
    olosses = []
    for k in range(steps):
        lo, g = cn.loss_grad(p.astype(np.float32), W, N, ins, outs, x, t, round_weights=True)
        olosses.append(lo)
        p, m1, m2 = cn.adam(p, g, m1, m2, k + 1, lr)
    print("losses", losses, olosses)
    assert np.allclose(losses, olosses, rtol=2e-5, atol=0)
    d_gpu, d_ora = net.params.astype(np.float64) - p0, p - p0
    assert _rel(d_gpu, d_ora) <= 2e-3


def _balanced_ensemble(lmw, N, seeds, precision=None):
    from lmweather_amd.corrector import EARTH, make_states
    e = lmw.Ensemble([N] * len(seeds), lmw.SphereParams(**EARTH), lmw.FP64 if precision is None else precision)
    make_states(e, seeds, N)
    return e


def test_truncation_round_trip(lmw):
    N, T = 21, 10
    e = _balanced_ensemble(lmw, N, [1, 2, 3])
    tr = e.get_truncated(T)
    C = cn.tri_size(T)
    deg, ordm = cn.tri_tables(T)
    for b in range(3):
        for f in range(3):
            s = e.get_spectral(b, f)
            want = s[ordm, deg]
            assert np.array_equal(tr[b, f, :, 0], want.real.astype(np.float32))
            assert np.array_equal(tr[b, f, :, 1], want.imag.astype(np.float32))
    e.set_truncated(T, tr)
    for b in range(3):
        for f in range(3):
            s = e.get_spectral(b, f)
            assert np.array_equal(s[ordm, deg].real, tr[b, f, :, 0].astype(np.float64))
            mask = np.ones_like(s, dtype=bool)
            mask[ordm, deg] = False
            assert np.all(s[mask] == 0)
    assert C == tr.shape[2]


def test_corrected_step_bookkeeping(lmw, corr):
    """Zero network: the corrected step is exactly the coarse sub-steps; threshold below every residual: exactly the
    fallback sub-steps; threshold between residuals: each member gets its own branch."""
    N, dt, seeds = 21, 2400.0, [4, 5, 6, 7]
    net = corr.Corrector(N, 16, seed=3)
    ref = _balanced_ensemble(lmw, N, seeds)
    x0 = ref.get_truncated(N)

    def fresh():
        e = lmw.Ensemble([N] * len(seeds), lmw.SphereParams(**corr.EARTH), lmw.FP64)
        e.set_truncated(N, x0)
        return e

    coarse, fine = fresh(), fresh()
    coarse.steps([dt / 2], [2])
    fine.steps([dt / 4], [4])
    xc, xf = coarse.get_truncated(N), fine.get_truncated(N)

    saved = net.params
    net.params = np.zeros_like(saved)
    e = fresh()
    used, res = corr.corrected_step(e, net, dt, 2, 4, np.inf)
    assert not used.any() and np.all(res == 0)
    assert np.array_equal(e.get_truncated(N), xc)
    assert e.diag(0, lmw._lib.DIAG_TIME) == dt

    net.params = saved
    e = fresh()
    used, res = corr.corrected_step(e, net, dt, 2, 4, -1.0)
    assert used.all()
    assert np.array_equal(e.get_truncated(N), xf)

    thr = float(np.median(res))
    e = fresh()
    used, res2 = corr.corrected_step(e, net, dt, 2, 4, thr)
    assert np.array_equal(res2, res)
    assert 0 < used.sum() < len(seeds)
    c, _ = net.forward(xc)
    got = e.get_truncated(N)
    for b in range(len(seeds)):
        if used[b]:
            assert np.array_equal(got[b], xf[b])
        else:
            want = xc[b].astype(np.float64) + c[b]
            for f in range(3):  # float32 snapshots of FP64 states: within float32 rounding of the field's scale
                scale = np.abs(want[f]).max()
                assert np.abs(got[b, f] - want[f]).max() <= 1e-6 * scale, (b, f)


def test_trained_corrector_gains_over_coarse_step(lmw, corr):
    """Held-out seeds, 1 day: the corrected step's error against the T170 model truncated to T42 is below that of
    the same coarse sub-steps without the network."""
    path = corr.DEFAULT_WEIGHTS
    if not os.path.exists(path):
        pytest.skip("no trained weights")
    net = corr.Corrector.load(path)
    out = corr.evaluate(net, list(range(100001, 100009)), days=1.0, leads_days=(0.5, 1), log=print)
    for lead in ("0.5", "1"):
        row = out["leads"][lead]
        print(lead, {k: (v["phi"], v["zeta"]) for k, v in row.items()})
        assert row["corrected"]["phi"] < row["coarse"]["phi"]
        assert row["corrected"]["zeta"] < row["coarse"]["zeta"]
    assert out["fallback_fraction"] <= 0.05
