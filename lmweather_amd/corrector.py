"""Learned coarse-step corrector (no reference counterpart): a spectral-space network that corrects a coarse
large-step T42 integration toward the T170 model truncated to T42, trained on the GPU with the library's own
forward, backward and Adam kernels (lmw_corrector_* in include/lmw.h).

The coarse step of ``dt`` (2400 s) is ``coarse_substeps`` RK4 steps of the plain spectral model (2 x 1200 s: one
RK4 step of 2400 s is unstable at T42, whose limit is near 1250 s for 8000 m depth and 40 m/s winds), then the
network's correction.  The plain model is the fallback (``fallback_substeps`` = 4 steps of 600 s) whenever the
network's output rms on a member exceeds a threshold set from the training distribution.

Pipeline (all on the GPU; ``python -m lmweather_amd.corrector --help``):

* ``generate``: T170 trajectories from seeded random balanced states (spectrum 1 <= n <= 21, max wind 40 m/s,
  depth 8000 m, no orography, no diffusion: the T42 test case's construction at T170), truncated to T42 every
  2400 s on the device; pairs (state, state 2400 s later); inputs = the coarse step of the first, targets = the
  second minus that.
* ``train``: per-degree scales from the data, then Adam in mini-batches over the device-resident training set.
* ``evaluate``: held-out seeds; T170 truth truncated to T42 against plain T42 at 600 s, the coarse step without the
  network, and the corrected step with fallback.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

from . import _lib as L
from .initial import random_balanced_state
from .sphere import Ensemble, Sphere

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "corrector_t42.npz")

EARTH = dict(radius=6.371e6, omega=7.292e-5, gravity=9.80616)


def tri_size(N: int) -> int:
    return (N + 1) * (N + 2) // 2


class Corrector:
    """Handle of the device network (lmw_corrector_*)."""

    def __init__(self, truncation: int = 42, width: int = 64, seed: int = 0):
        h = C.c_void_p()
        L.call("lmw_corrector_create", int(truncation), int(width), int(seed), C.byref(h))
        self._h = h
        self.N, self.width = int(truncation), int(width)
        n = C.c_size_t()
        L.call("lmw_corrector_dims", self._h, None, None, C.byref(n))
        self.nparams = n.value
        self.C = tri_size(self.N)
        self.in_scale = np.ones(3 * (self.N + 1), dtype=np.float32)
        self.out_scale = np.ones(3 * (self.N + 1), dtype=np.float32)
        self.threshold = math.inf
        self.coarse_substeps, self.fallback_substeps, self.dt = 2, 4, 2400.0
        self.meta: dict = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                L.lib().lmw_corrector_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        L.call("lmw_corrector_get_params", self._h, L.fptr(out), out.size)
        return out

    @params.setter
    def params(self, p) -> None:
        a = L.as_f32(p).ravel()
        L.call("lmw_corrector_set_params", self._h, L.fptr(a), a.size)

    def set_scales(self, in_scale, out_scale) -> None:
        a, b = L.as_f32(in_scale).ravel(), L.as_f32(out_scale).ravel()
        L.call("lmw_corrector_set_scales", self._h, L.fptr(a), L.fptr(b), a.size)
        self.in_scale, self.out_scale = a.copy(), b.copy()

    def _check(self, x) -> np.ndarray:
        a = L.as_f32(x)
        if a.ndim != 4 or a.shape[1:] != (3, self.C, 2):
            raise ValueError(f"expected (samples, 3, {self.C}, 2) spectra")
        return a

    def loss_grad(self, x, target, grad: bool = True):
        a, t = self._check(x), self._check(target)
        loss = C.c_double()
        g = np.empty(self.nparams, dtype=np.float32) if grad else None
        L.call("lmw_corrector_loss_grad", self._h, L.fptr(a), L.fptr(t), a.shape[0], C.byref(loss),
               L.fptr(g) if grad else None)
        return (loss.value, g) if grad else loss.value

    def forward(self, x):
        a = self._check(x)
        out = np.empty_like(a)
        res = np.empty(a.shape[0])
        L.call("lmw_corrector_forward", self._h, L.fptr(a), a.shape[0], L.fptr(out), L.dptr(res))
        return out, res

    def set_dataset(self, x, target) -> None:
        a, t = self._check(x), self._check(target)
        L.call("lmw_corrector_set_dataset", self._h, L.fptr(a), L.fptr(t), a.shape[0])

    def train(self, epochs: int, batch: int, lr: float, seed: int = 0) -> np.ndarray:
        losses = np.zeros(max(1, epochs))
        L.call("lmw_corrector_train", self._h, int(epochs), int(batch), float(lr), int(seed), L.dptr(losses))
        return losses[:epochs]

    # ------------------------------------------------------------------ persistence
    def save(self, path: str) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, params=self.params, in_scale=self.in_scale, out_scale=self.out_scale,
                 dims=np.array([self.N, self.width, self.coarse_substeps, self.fallback_substeps]),
                 dt=np.array([self.dt]), threshold=np.array([self.threshold]), meta=np.array(json.dumps(self.meta)))

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS) -> "Corrector":
        z = np.load(path, allow_pickle=False)
        N, W, nc, nf = (int(v) for v in z["dims"])
        c = cls(N, W)
        c.params = z["params"]
        c.set_scales(z["in_scale"], z["out_scale"])
        c.coarse_substeps, c.fallback_substeps = nc, nf
        c.dt, c.threshold = float(z["dt"][0]), float(z["threshold"][0])
        c.meta = json.loads(str(z["meta"]))
        return c


def corrected_step(model, corrector: Corrector, dt: float | None = None, coarse_substeps: int | None = None,
                   fallback_substeps: int | None = None, threshold: float | None = None):
    """One corrected coarse step of a Sphere or a single-truncation Ensemble; returns (used_fallback, residual)
    per member."""
    dt = corrector.dt if dt is None else dt
    nc = corrector.coarse_substeps if coarse_substeps is None else coarse_substeps
    nf = corrector.fallback_substeps if fallback_substeps is None else fallback_substeps
    thr = corrector.threshold if threshold is None else threshold
    B = len(model.truncations) if isinstance(model, Ensemble) else 1
    used = np.zeros(B, dtype=np.int32)
    res = np.zeros(B)
    fn = "lmw_ensemble_corrected_step" if isinstance(model, Ensemble) else "lmw_sphere_corrected_step"
    L.call(fn, model._h, corrector._h, float(dt), int(nc), int(nf), float(thr), used.ctypes.data_as(L._ip), L.dptr(res))
    return used.astype(bool), res


# ---------------------------------------------------------------------------------------------- data and metrics
def make_states(ens: Ensemble, seeds, N: int, nmax: int = 21) -> None:
    """Member i of `ens` (truncation N) <- the random balanced state of seeds[i]."""
    base = Sphere(N, L.SphereParams(**EARTH))
    for i, s in enumerate(seeds):
        h, u, v = random_balanced_state(base, nmax, int(s))
        ens.set_state_grid(i, h, u, v)


def l2_weights(N: int) -> np.ndarray:
    """Weights of |c|^2 that give the L2 norm of a real field from its m >= 0 coefficients (2 for m > 0)."""
    w = np.empty(tri_size(N))
    c = 0
    for m in range(N + 1):
        w[c:c + N + 1 - m] = 1.0 if m == 0 else 2.0
        c += N + 1 - m
    return w


def field_errors(x, truth, N: int) -> np.ndarray:
    """Relative L2 errors (members, 3) of zeta, delta and phi' (phi without its global mean) of spectra (B, 3, C, 2);
    delta is relative to the norm of zeta (it is small in balanced flow)."""
    w = l2_weights(N)
    w0 = w.copy()
    w0[0] = 0.0
    d = np.asarray(x, dtype=np.float64) - np.asarray(truth, dtype=np.float64)
    t = np.asarray(truth, dtype=np.float64)
    sq = lambda a, ww: np.sqrt(((a * a).sum(-1) * ww).sum(-1))  # noqa: E731
    dz, dd, dp = sq(d[:, 0], w0), sq(d[:, 1], w0), sq(d[:, 2], w0)
    tz, tp = sq(t[:, 0], w0), sq(t[:, 2], w0)
    return np.stack([dz / tz, dd / tz, dp / tp], axis=1)


def generate(seeds, days: float = 10.0, T_hi: int = 170, dt_hi: float = 150.0, T_lo: int = 42, dt: float = 2400.0,
             coarse_substeps: int = 2, pair_stride: int = 4, chunk: int = 256, precision: int = L.BF16X3, log=print):
    """Training pairs from T170 trajectories: returns (inputs, targets), float32 (n, 3, C42, 2)."""
    per = int(round(dt / dt_hi))
    if abs(per * dt_hi - dt) > 1e-9:
        raise ValueError("dt must be a multiple of dt_hi")
    nint = int(round(days * 86400.0 / dt))
    starts = list(range(0, nint, pair_stride))
    needed = sorted(set(starts) | {k + 1 for k in starts})
    xs, ts = [], []
    t0 = time.time()
    for c0 in range(0, len(seeds), chunk):
        sd = seeds[c0:c0 + chunk]
        hi = Ensemble([T_hi] * len(sd), L.SphereParams(**EARTH), precision)
        make_states(hi, sd, T_hi)
        snaps = {}
        for k in range(nint + 1):
            if k in needed:
                snaps[k] = hi.get_truncated(T_lo)
            if k < nint and k + 1 <= needed[-1]:
                hi.steps([dt_hi], [per])
        del hi
        x0 = np.concatenate([snaps[k] for k in starts])
        x1 = np.concatenate([snaps[k + 1] for k in starts])
        xc = coarse_steps(x0, T_lo, dt, coarse_substeps)
        xs.append(xc)
        ts.append(x1 - xc)
        log(f"generate: seeds {c0}..{c0 + len(sd) - 1}: {len(starts) * len(sd)} pairs, {time.time() - t0:.1f} s")
    return np.concatenate(xs), np.concatenate(ts)


def coarse_steps(x0, T: int, dt: float, substeps: int, chunk: int = 4096) -> np.ndarray:
    """The coarse step (substeps FP64 RK4 steps of dt/substeps at truncation T) of every state in x0."""
    out = np.empty_like(x0)
    for a in range(0, x0.shape[0], chunk):
        b = min(a + chunk, x0.shape[0])
        e = Ensemble([T] * (b - a), L.SphereParams(**EARTH), L.FP64)
        e.set_truncated(T, x0[a:b])
        e.steps([dt / substeps], [substeps])
        out[a:b] = e.get_truncated(T)
    return out


def degree_scales(data, N: int) -> np.ndarray:
    """Per-field, per-degree rms over samples and orders of spectra (n, 3, C, 2): float32 (3, N+1), floored."""
    deg = np.concatenate([np.arange(m, N + 1) for m in range(N + 1)])
    sq = np.zeros((3, tri_size(N)))
    for a in range(0, len(data), 4096):  # chunks: no float64 copy of the whole set
        d = np.asarray(data[a:a + 4096], dtype=np.float64)
        sq += (d * d).sum(-1).sum(0)
    sq /= len(data)  # (3, C)
    out = np.zeros((3, N + 1))
    cnt = np.bincount(deg, minlength=N + 1)
    for f in range(3):
        out[f] = np.sqrt(np.bincount(deg, weights=sq[f], minlength=N + 1) / cnt)
    for f in range(3):
        pos = out[f][out[f] > 0]
        floor = pos.max() * 1e-6 if pos.size else 1.0
        out[f] = np.where(out[f] > floor, out[f], floor if pos.size else 1.0)
    return out.astype(np.float32)


def trained_mask(N: int) -> np.ndarray:
    """(C, 3, 2) bool: the entries the network trains and corrects (all but the (0, 0) row and the imaginary parts
    at m = 0)."""
    m = np.ones((tri_size(N), 3, 2), dtype=bool)
    m[0] = False
    m[:N + 1, :, 1] = False  # the m = 0 block is the first N + 1 packed entries
    return m


def zero_correction_loss(t, out_scale, N: int) -> float:
    """The training loss of a network that outputs zero: mean of (target / out_scale)^2 over trained entries."""
    deg = np.concatenate([np.arange(m, N + 1) for m in range(N + 1)])
    sc = np.asarray(out_scale, dtype=np.float64).reshape(3, N + 1)[:, deg]  # (3, C)
    v = np.asarray(t, dtype=np.float64) / sc[None, :, :, None]
    mask = trained_mask(N).transpose(1, 0, 2)[None]  # (1, 3, C, 2)
    return float((v * v * mask).sum() / (mask.sum() * v.shape[0]))


def fit(x, t, width: int = 64, epochs: int = 12, batch: int = 32, lr: float = 1e-3, seed: int = 0, holdout: float = 0.05,
        log=print) -> Corrector:
    """Train a corrector on inputs x and targets t (n, 3, C, 2); a fraction `holdout` of the pairs is kept out to
    set the fallback threshold (the 99.9th percentile of the residual on unseen pairs, times 1.5)."""
    N = int(round((math.sqrt(8 * x.shape[2] + 1) - 3) / 2))
    n = x.shape[0]
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    nv = max(1, int(n * holdout))
    val, tr = perm[:nv], perm[nv:]
    net = Corrector(N, width, seed)
    net.set_scales(degree_scales(x[tr], N), degree_scales(t[tr], N))
    net.set_dataset(x[tr], t[tr])
    base = zero_correction_loss(t[val], net.out_scale, N)
    log(f"train: {tr.size} pairs, {nv} held out; loss of a zero correction on held-out pairs {base:.4e}")
    history = []
    t0 = time.time()
    for e in range(epochs):
        rate = lr * 0.5 * (1 + math.cos(math.pi * e / max(1, epochs)))  # cosine schedule
        lt = float(net.train(1, batch, rate, seed + e)[0])
        lv = float(np.mean([net.loss_grad(x[val[a:a + 256]], t[val[a:a + 256]], grad=False) for a in range(0, nv, 256)]))
        history.append((lt, lv))
        log(f"train: epoch {e + 1}/{epochs} lr {rate:.2e} loss {lt:.4e} held-out {lv:.4e} ({time.time() - t0:.1f} s)")
    res = np.concatenate([net.forward(x[val[a:a + 256]])[1] for a in range(0, nv, 256)])
    net.threshold = float(np.quantile(res, 0.999) * 1.5)
    secs = time.time() - t0
    steps = epochs * (tr.size // batch)
    net.meta.update(train_pairs=int(tr.size), holdout_pairs=int(nv), epochs=epochs, batch=batch, lr=lr, width=width,
                    zero_correction_loss=base, history=history, train_seconds=secs, adam_steps=steps,
                    rows_per_second=steps * batch * tri_size(N) / max(secs, 1e-9), threshold=net.threshold)
    return net


def evaluate(net: Corrector, seeds, days: float = 10.0, T_hi: int = 170, dt_hi: float = 150.0, dt_plain: float = 600.0,
             precision: int = L.BF16X3, leads_days=(0.5, 1, 2, 5, 10), log=print) -> dict:
    """Errors against the T170 model truncated to T42 for plain T42 (dt_plain), the coarse step without the
    network, and the corrected step with fallback; mean over members at the given leads, plus wall time."""
    N, dt = net.N, net.dt
    per = int(round(dt / dt_hi))
    nint = int(round(days * 86400.0 / dt))
    B = len(seeds)
    hi = Ensemble([T_hi] * B, L.SphereParams(**EARTH), precision)
    make_states(hi, seeds, T_hi)
    truth = [hi.get_truncated(N)]
    for _ in range(nint):
        hi.steps([dt_hi], [per])
        truth.append(hi.get_truncated(N))
    del hi
    log(f"evaluate: truth for {B} seeds, {days} days")
    runs = {}
    names = ("plain", "coarse", "corrected")
    ens = {k: Ensemble([N] * B, L.SphereParams(**EARTH), L.FP64) for k in names}
    for e in ens.values():
        e.set_truncated(N, truth[0])
    errs = {k: [np.zeros((B, 3))] for k in names}
    wall = dict.fromkeys(names, 0.0)
    nfall = 0
    nplain = int(round(dt / dt_plain))
    for k in range(1, nint + 1):
        t0 = time.perf_counter()
        ens["plain"].steps([dt_plain], [nplain])
        ens["plain"].get_truncated(N)
        wall["plain"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        ens["coarse"].steps([dt / net.coarse_substeps], [net.coarse_substeps])
        ens["coarse"].get_truncated(N)
        wall["coarse"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        used, _ = corrected_step(ens["corrected"], net)
        wall["corrected"] += time.perf_counter() - t0
        nfall += int(used.sum())
        for name in names:
            errs[name].append(field_errors(ens[name].get_truncated(N), truth[k], N))
    out = {"members": B, "days": days, "fallback_fraction": nfall / (B * nint), "wall_seconds": wall, "leads": {}}
    for ld in leads_days:
        k = int(round(ld * 86400.0 / dt))
        if k > nint:
            continue
        row = {}
        for name in names:
            e = errs[name][k]
            bad = ~np.isfinite(e).all(axis=1)
            row[name] = {"zeta": float(np.nanmean(e[:, 0])), "delta": float(np.nanmean(e[:, 1])),
                         "phi": float(np.nanmean(e[:, 2])), "nonfinite_members": int(bad.sum())}
        out["leads"][str(ld)] = row
        log(f"evaluate: {ld} d: " + ", ".join(f"{n} phi {row[n]['phi']:.3e} zeta {row[n]['zeta']:.3e}" for n in names))
    return out


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description="learned coarse-step corrector: generate, train, evaluate (GPU)")
    ap.add_argument("--train-seeds", type=int, default=2000)
    ap.add_argument("--eval-seeds", type=int, default=256)
    ap.add_argument("--days", type=float, default=10.0)
    ap.add_argument("--eval-days", type=float, default=10.0)
    ap.add_argument("--pair-stride", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--width", type=int, default=64)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--report", default=None, help="JSON file for the training and evaluation summary")
    ap.add_argument("--skip-train", action="store_true", help="evaluate the weights in --out")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    report = {}
    if not a.skip_train:
        x, t = generate(list(range(1, a.train_seeds + 1)), days=a.days, pair_stride=a.pair_stride, log=log)
        net = fit(x, t, width=a.width, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
        net.meta.update(train_seeds=a.train_seeds, days=a.days, pair_stride=a.pair_stride)
        del x, t
        net.save(a.out)
        log(f"saved {a.out} (threshold {net.threshold:.4g})")
        report["train"] = net.meta
    else:
        net = Corrector.load(a.out)
    if a.eval_seeds > 0:
        report["evaluate"] = evaluate(net, list(range(100001, 100001 + a.eval_seeds)), days=a.eval_days, log=log)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
