"""Benchmark: ensemble member-days/s at T170 on one MI355X, mixed precision vs FP64.

Workload (one "step" = one RK4 step of the whole batch): ``--members`` members at T170 (512 x 256 Gaussian grid,
a = 6.371e6 m, dt = 150 s) -- the T170 share of the 256-member uneven ensemble -- each the T170 random balanced state
(energy up to n = 85, seed 1) over 8 seeded Gaussian mountains, with its vorticity coefficients perturbed by 1e-3
relative noise (seed 1000 + member).  Inputs are resident in HBM before the timed region.

value = whole-job member-days/s of the batch with the Legendre contractions in ``--precision`` (default BF16x6:
FP32 operands split into three BF16 parts, six MFMA products, FP32 accumulation; the FFTs and grid products between
the two contractions run in FP32 in the mixed modes); the same batch in FP64 and one FP64 member ("plain") are timed
beside it.  The accuracy of the mixed precision is measured on the timed batch itself: every member's rel-L2 of h, u,
v against the same member of the FP64 batch after ``--check-days`` simulated days (worst and mean over the members),
next to an FP64 twin batch started from a further 1e-7 relative vorticity perturbation (the dynamics' own error
growth).  Other workloads: ``--workload ensemble`` (C3), ``plane`` (C4), ``setup``, ``c1`` (one T42 member, the
reference's shape) and ``accuracy`` (C2 single instance in every mode to 1 and ``--check-days`` days).  The roofline is that of
the dominant kernel, measured with HIP events around every launch on the library's stream in a second pass of as
many steps right after the timed region.  The reference CPU code
(oracle/_ref, one core) is timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

A, OM, G = 6.371e6, 7.292e-5, 9.80616
N, DT = 170, 150.0
PEAKS = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16x3": 2.5e15, "bf16x6": 2.5e15}
MFMA_PRODUCTS = {"fp64": 1.0, "fp32": 1.0, "bf16x3": 3.0, "bf16x6": 6.0}  # MFMA products per multiply-add
PREC_NAMES = ("fp64", "fp32", "bf16x3", "bf16x6")
DTYPES = {"fp64": "f64",
          "fp32": "f32 Legendre and f32 FFT/grid products; f64 state/RK4/tables",
          "bf16x3": "bf16x3 Legendre (f32 acc) and f32 FFT/grid products; f64 state/RK4/tables",
          "bf16x6": "bf16x6 Legendre (f32 acc) and f32 FFT/grid products; f64 state/RK4/tables"}
HBM_PEAK = 8.0e12


def measured_traffic(name: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this exact
    workload (profiles/<name>_traffic.txt, copy-calibrated), or None when there is no summary for it."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1].split("::")[-1] == kernel:
            return float(f[2])
    return None


def initial_state(lmw):
    from lmweather_amd.initial import gaussian_orography, random_balanced_state
    base = lmw.Sphere(N, lmw.SphereParams(A, OM, G))
    h, u, v = random_balanced_state(base, 85, 1)
    hs = gaussian_orography(base, 8, 2)
    return h - hs, u, v, hs


def build_batch(lmw, members, precision, state):
    from lmweather_amd.initial import perturb_vorticity
    h, u, v, hs = state
    ens = lmw.Ensemble([N] * members, lmw.SphereParams(A, OM, G), precision)
    for i in range(members):
        ens.set_orography_grid(i, hs)
        ens.set_state_grid(i, h, u, v)
        ens.set_spectral(i, 0, perturb_vorticity(ens.get_spectral(i, 0), 1e-3, 1000 + i))
    ens.autotune()  # Legendre tile shapes: timed on this batch, or taken from the tile cache
    return ens


def kernel_costs(J, K, S, M1, members, mixed, orography_planes=1):
    """Algorithmic work per launch (one RK4 stage of the batch).  Legendre flops as executed with the equatorial
    folding (half of the reference's 24 J S and 28 J S per member); for split BF16 the MFMA executes three products
    per multiply-add.  Grid kernel HBM bytes: the four Fourier coefficients of every (latitude, m) in, the analysis
    operands out -- in the mixed modes one FP32 record of ten complex values (five weighted coefficients at each
    latitude of the pair) per (m, latitude pair), in FP64 sixteen complex operands (both parities' north/south
    combinations) -- per member, and each distinct orography grid once (members with the same orography share one
    device plane; the benchmark's members all share one)."""
    Jh = J // 2
    cb = 8 if mixed else 16  # bytes per stored complex
    per_pair = 10 if mixed else 16  # stored complex values per (m, latitude pair)
    synth_flops = 24.0 * Jh * S * members
    anal_flops = 28.0 * Jh * S * members
    grid_bytes = (4 * J * M1 * cb + M1 * Jh * per_pair * cb) * members + J * K * 8 * orography_planes
    return synth_flops, grid_bytes, anal_flops


def timed_batch(ens, steps, warmup):
    """Wall time of `steps` RK4 steps on the production path (replays of a HIP graph of 8 steps, then of a one-step graph
    for the rest; synchronous), then the per-kernel times of as many steps again with HIP events around every kernel on
    the library's stream (the events add a few microseconds between kernels, so they stay out of the timed region)."""
    ens.steps([DT], [warmup])
    t0 = time.perf_counter()
    ens.steps([DT], [steps])
    wall = time.perf_counter() - t0
    ms = ens.steps_timed([DT], [steps])[0]
    return wall, ms


def _cpu_member(args):
    """One reference member on one core (run in a worker process): wall seconds of `steps` RK4 steps."""
    state, steps, kind = args
    from oracle import pyapi
    h, u, v, hs = state
    s = pyapi.OracleSphere(N, A, OM, G, kind=kind)
    s.set_orography_grid(hs)
    s.set_state_grid(h, u, v)
    s.step(DT)  # warm caches
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(DT)
    return time.perf_counter() - t0


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    """Cores this process may use, at most 16 (the share of one GPU's host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(state, steps):
    """The reference's own code (oracle/_ref, else the oracle restatement) on the host cores: `steps` RK4 steps of
    one T170 member on one core, then P independent members on P cores at once, one worker process each (the
    reference is single-threaded; an ensemble spreads over cores).  value = the P-core ensemble rate."""
    import multiprocessing as mp
    from oracle import pyapi
    kind = "reference" if pyapi.available("ref") else "port"
    lib = "ref" if kind == "reference" else "oracle"
    wall1 = _cpu_member((state, steps, lib))
    one = DT * steps / 86400.0 / wall1
    P = host_cores()
    psteps = max(8, steps // 4)
    ctx = mp.get_context("spawn")  # fresh interpreters: nothing of this process's GPU context is inherited
    t0 = time.perf_counter()
    with ctx.Pool(P) as pool:
        walls = pool.map(_cpu_member, [(state, psteps, lib)] * P)
    wall_p = time.perf_counter() - t0
    agg = P * DT * psteps / 86400.0 / max(walls)
    return {"value": agg, "unit": "member-days/s", "cores": P, "kind": kind, "cpu": cpu_model(),
            "sample": f"{P} independent T170 members, {psteps} RK4 steps (dt {DT:g} s) each, one per core in "
                      f"{P} worker processes, slowest {max(walls):.2f} s ({wall_p:.2f} s with process start-up)",
            "single_core": {"value": one, "cores": 1, "ms_per_step": 1e3 * wall1 / steps,
                            "sample": f"{steps} RK4 steps of one T170 member, single thread, {wall1:.2f} s"}}


def batch_drifts(lmw, ens, members):
    """(energy, mean height) of every member: the start of the conservation drifts."""
    return [(ens.diag(i, lmw._lib.DIAG_ENERGY), ens.diag(i, lmw._lib.DIAG_MEAN_HEIGHT)) for i in range(members)]


def batch_accuracy(lmw, ens, e64, twin, done, members, days, start):
    """Accuracy of the timed batch itself: every member of the mixed batch `ens` and of the FP64 twin batch `twin`
    (the FP64 batch with a further 1e-7 relative vorticity perturbation: the dynamics' own error growth) against the
    same member of the FP64 batch `e64`, all advanced to `days` simulated days (done[k] = steps batch k has taken in
    the timing passes).  Worst and mean over the members of rel-L2 of h, u, v; worst energy and mass drift."""
    n = int(round(days * 86400 / DT))
    for e, d in zip((ens, e64, twin), done):
        if n > d:
            e.steps([DT], [n - d])
    out = {"days": days, "steps": n, "members": members, "measured_on": "the timed batch (all members)"}
    for name, e in (("mixed", ens), ("twin", twin)):
        for f, fld in enumerate("huv"):
            errs = np.array([np.linalg.norm(e.grid(i, f) - e64.grid(i, f)) / np.linalg.norm(e64.grid(i, f))
                             for i in range(members)])
            out[f"{name}_rel_l2_{fld}_worst"] = float(errs.max())
            out[f"{name}_rel_l2_{fld}_mean"] = float(errs.mean())
    for name, e, k in (("fp64", e64, 1), ("mixed", ens, 0)):
        de = [abs(e.diag(i, lmw._lib.DIAG_ENERGY) / start[k][i][0] - 1.0) for i in range(members)]
        dm = [abs(e.diag(i, lmw._lib.DIAG_MEAN_HEIGHT) / start[k][i][1] - 1.0) for i in range(members)]
        out[f"{name}_energy_drift_worst"] = float(max(de))
        out[f"{name}_mass_drift_worst"] = float(max(dm))
    return out


def single_accuracy(lmw, state, days_list):
    """Config C2 as a single instance: one T170 member (the C2 state over the mountains, no perturbation) stepped in
    FP64, FP32, BF16x3 and BF16x6, and an FP64 twin from a 1e-7 relative vorticity perturbation, to each of
    `days_list`; rel-L2 of h, u, v against FP64 and the energy and mass drift of each."""
    from lmweather_amd.initial import perturb_vorticity
    h, u, v, hs = state
    runs = {}
    for name, pr, twin in (("fp64", lmw.FP64, False), ("fp32", lmw.FP32, False), ("bf16x3", lmw.BF16X3, False),
                           ("bf16x6", lmw.BF16X6, False), ("fp64_twin", lmw.FP64, True)):
        s = lmw.Sphere(N, lmw.SphereParams(A, OM, G))
        s.set_orography_grid(hs)
        s.set_state_grid(h, u, v)
        if twin:
            s.set_spectral(0, perturb_vorticity(s.get_spectral(0), 1e-7, 77))
        if pr != lmw.FP64:
            s.set_precision(pr)
        runs[name] = (s, s.energy(), s.mean_height())
    out, done = {}, 0
    for days in days_list:
        n = int(round(days * 86400 / DT))
        t0 = time.perf_counter()
        for s, _, _ in runs.values():
            s.steps(DT, n - done)
        done = n
        print(f"accuracy: day {days:g} reached in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        ref = runs["fp64"][0]
        rg = [ref.grid_height(), ref.grid_u(), ref.grid_v()]
        rec = {}
        for name, (s, e0, m0) in runs.items():
            r = {"energy_drift": s.energy() / e0 - 1.0, "mass_drift": s.mean_height() / m0 - 1.0}
            if name != "fp64":
                for f, g in zip("huv", (s.grid_height(), s.grid_u(), s.grid_v())):
                    r[f"rel_l2_{f}"] = float(np.linalg.norm(g - rg["huv".index(f)]) / np.linalg.norm(rg["huv".index(f)]))
            rec[name] = r
        out[f"day_{days:g}"] = rec
    return out


def c1_cpu_reference(state, seconds, dt):
    """CPU baseline leg of config C1: the reference's own code (oracle/_ref, else the oracle) on one core runs the
    same run(seconds, dt) from the same grids; returns its timing, its final height grid (the checker of the GPU
    result) and its drifts."""
    from oracle import pyapi
    kind = "reference" if pyapi.available("ref") else "port"
    h, u, v = state
    s = pyapi.OracleSphere(42, A, OM, G, kind="ref" if kind == "reference" else "oracle")
    s.set_state_grid(h, u, v)
    e0, q0, m0 = s.energy(), s.potential_enstrophy(), s.mean_height()
    t0 = time.perf_counter()
    s.run(seconds, dt)
    wall = time.perf_counter() - t0
    steps = int(round(seconds / dt))
    return {"value": seconds / 86400.0 / wall, "unit": "member-days/s", "cores": 1, "kind": kind, "cpu": cpu_model(),
            "sample": f"the whole run: {steps} RK4 steps (dt {dt:g} s) of one T42 member, single thread, {wall:.2f} s",
            "ms_per_step": 1e3 * wall / steps}, s.grid_height(), {
        "energy_drift": s.energy() / e0 - 1.0, "enstrophy_drift": s.potential_enstrophy() / q0 - 1.0,
        "mass_drift": s.mean_height() / m0 - 1.0}


def c1_bench(lmw, args) -> dict:
    """Config C1: one T42 member (128 x 64 Gaussian grid, a = 6.371e6 m), the balanced random flow n <= 21 (seed 1,
    max |u| 40 m/s, depth 8000 m), FP64, run(5 days, 600 s) = 720 RK4 steps with the reference's checks every 64
    steps -- the reference's own shape, a single instance.  value = member-days/s of the GPU run; beside it the
    reference on one core over the same run, the GPU's rel-L2(h) against it at day 5 and the drifts of both."""
    from lmweather_amd.initial import random_balanced_state
    seconds, dt = 5 * 86400.0, 600.0
    s = lmw.Sphere(42, lmw.SphereParams(A, OM, G))
    h, u, v = random_balanced_state(s, 21, 1)
    s.set_state_grid(h, u, v)
    s.steps(dt, max(args.warmup, 1))  # graph capture and first launches
    s.set_state_grid(h, u, v)
    e0, q0, m0 = s.energy(), s.potential_enstrophy(), s.mean_height()
    t0 = time.perf_counter()
    s.run(seconds, dt)
    wall = time.perf_counter() - t0
    steps = int(round(seconds / dt))
    out = {
        "metric": "member-days/s, single T42 instance (FP64)", "value": seconds / 86400.0 / wall,
        "unit": "member-days/s", "n_gpus": 1, "steps": steps, "warmup": max(args.warmup, 1),
        "ms_per_step": 1e3 * wall / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "T42 single instance (128x64 Gaussian grid), balanced random flow n<=21 seed 1, "
                               "run(5 days, dt 600 s) with the checks every 64 steps", "truncation": 42, "dt_s": dt},
        "drift": {"energy": s.energy() / e0 - 1.0, "enstrophy": s.potential_enstrophy() / q0 - 1.0,
                  "mass": s.mean_height() / m0 - 1.0},
    }
    if not args.no_cpu:
        print("cpu baseline: the reference's 720 steps", file=sys.stderr, flush=True)
        out["cpu_baseline"], href, out["cpu_drift"] = c1_cpu_reference((h, u, v), seconds, dt)
        hg = s.grid_height()
        out["rel_l2_h_vs_reference_day5"] = float(np.linalg.norm(hg - href) / np.linalg.norm(href))
        out["speedup_vs_cpu_core"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def accuracy_bench(lmw, args) -> dict:
    """Config C2 accuracy: one T170 member in every precision mode and an FP64 twin, to 1 and 10 days."""
    state = initial_state(lmw)
    days = [1.0, args.check_days] if args.check_days > 1.0 else [args.check_days]
    t0 = time.perf_counter()
    res = single_accuracy(lmw, state, days)
    return {"metric": "rel-L2 of h, u, v against FP64, single T170 instance", "value": res[f"day_{days[-1]:g}"]["bf16x6"]["rel_l2_h"],
            "unit": "rel-L2(h) of BF16x6 at the last day", "n_gpus": 1, "steps": int(round(days[-1] * 86400 / DT)),
            "warmup": 0, "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64 reference; f32, bf16x3, bf16x6 Legendre with f32 FFT/grid products", "data": "synthetic",
            "config": {"workload": "T170 single instance (512x256 Gaussian grid, dt 150 s), balanced random flow n<=85 "
                                   "over 8 Gaussian mountains (C2)", "days": days},
            "wall_s": time.perf_counter() - t0, "accuracy": res}


def diagnostics_bench(lmw, args) -> dict:
    """Diagnostics and the checks of run() on one T170 member (the C2 state): each scalar diagnostic through the
    C ABI (one device synthesis of the grids, one compensated reduction, the scalar copied to the host), and run()
    against steps() over the same 640 steps (run checks finiteness and the stability limit every 64 steps)."""
    h, u, v, hs = initial_state(lmw)
    s = lmw.Sphere(N, lmw.SphereParams(A, OM, G))
    s.set_orography_grid(hs)
    s.set_state_grid(h, u, v)
    s.energy()
    out = {}
    for name in ("energy", "potential_enstrophy", "max_wind", "height_error_l2", "mean_height"):
        fn = getattr(s, name)
        fn()
        t0 = time.perf_counter()
        for _ in range(50):
            fn()
        out[f"{name}_ms"] = 1e3 * (time.perf_counter() - t0) / 50
    t0 = time.perf_counter()
    for _ in range(20):
        s.stable_dt()
    out["stable_dt_ms"] = 1e3 * (time.perf_counter() - t0) / 20
    n = 640
    s.steps(DT, 64)
    walls = {}
    for mode in ("steps", "run", "steps", "run"):
        s.set_state_grid(h, u, v)
        t0 = time.perf_counter()
        if mode == "steps":
            s.steps(DT, n)
        else:
            s.run(n * DT, DT)
        w = time.perf_counter() - t0
        walls[mode] = min(walls.get(mode, w), w)
    return {"metric": "diagnostics and run() checks, single T170 instance", "value": out["energy_ms"], "unit": "ms per energy()",
            "n_gpus": 1, "steps": n, "warmup": 64, "ms_per_step": 1e3 * walls["steps"] / n, "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "T170 single instance (512x256 Gaussian grid), C2 state over 8 Gaussian mountains"},
            "diagnostics": out, "steps_s": walls["steps"], "run_s": walls["run"],
            "run_overhead_frac": walls["run"] / walls["steps"] - 1.0}


def corrector_bench(lmw, args) -> dict:
    """The learned corrector's kernels (T42, width 64): training throughput (shuffled mini-batches of 32 samples from a
    device-resident set of 2048 synthetic samples, forward + backward + Adam per mini-batch) and inference throughput
    (the forward pass of 256 samples from host arrays), in rows (sample x coefficient) per second.  value = training
    rows/s."""
    from lmweather_amd.corrector import Corrector
    rng = np.random.default_rng(5)
    net = Corrector(42, 64, seed=1)
    x = rng.standard_normal((2048, 3, net.C, 2)).astype(np.float32)
    t = (0.1 * rng.standard_normal((2048, 3, net.C, 2))).astype(np.float32)
    net.set_dataset(x, t)
    net.train(1, 32, 1e-3, seed=1)  # warm-up: buffers, first launches
    epochs = 3
    t0 = time.perf_counter()
    losses = net.train(epochs, 32, 1e-3, seed=2)
    train_s = time.perf_counter() - t0
    rows_train = epochs * 2048 * net.C / train_s
    xi = x[:256]
    net.forward(xi)
    t0 = time.perf_counter()
    for _ in range(5):
        net.forward(xi)
    fwd_s = (time.perf_counter() - t0) / 5
    return {"metric": "learned corrector training rows/s (T42, width 64)", "value": rows_train, "unit": "rows/s",
            "n_gpus": 1, "steps": epochs * 2048 // 32, "warmup": 2048 // 32, "ms_per_step": 1e3 * train_s / (epochs * 2048 // 32),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16 weights, activations split into three bf16 parts on MFMA (f32-exact products), f32 accumulation",
            "data": "synthetic", "config": {"workload": "corrector T42 width 64: 2048 synthetic samples, batch 32, 3 epochs; "
                                                        "inference 256 samples"},
            "inference_rows_per_s": 256 * net.C / fwd_s, "inference_ms_256_samples": 1e3 * fwd_s,
            "final_epoch_loss": float(losses[-1])}


def plane_jet_state(nx, ny, lx, ly, g, H, f0, beta, U=1.0, Lw=1.0e5):
    """Geostrophically balanced sech^2 jet: u = U sech^2(s) at u points, h = H - (f0 U L / g) tanh s
    - (beta U L^2 / g)(s tanh s - ln cosh s) at centres, s = (y - ly/2) / L (config C4)."""
    dy = ly / ny
    y = (np.arange(ny) + 0.5) * dy
    s = (y - 0.5 * ly) / Lw
    u = np.repeat((U / np.cosh(s) ** 2)[:, None], nx, axis=1)
    hrow = H - (f0 * U * Lw / g) * np.tanh(s) - (beta * U * Lw * Lw / g) * (s * np.tanh(s) - np.log(np.cosh(s)))
    h = np.repeat(hrow[:, None], nx, axis=1)
    return h, u


def plane_bench(lmw, args) -> dict:
    """Config C4: 1024 x 512 beta-plane channel (dx = dy = 10 km, f0 = 1e-4, beta = 1.6e-11, H = 1000 m), a batch of
    members each the sech^2 jet plus white height noise of 1e-3 m (seed 3 + member), RK4 at dt = stable_dt(0.5)."""
    from lmweather_amd.initial import normals
    nx, ny, lx, ly, g, H, f0, beta = 1024, 512, 1.024e7, 5.12e6, 9.81, 1000.0, 1e-4, 1.6e-11
    prec = lmw.FP64 if args.precision == "fp64" else lmw.FP32
    p = lmw.PlaneParams(nx, ny, lx, ly, g, H, f0, beta)
    pl = lmw.Plane(p, args.members, prec)
    if args.plane_tile is not None:
        pl.set_tile(args.plane_tile)
    h, u = plane_jet_state(nx, ny, lx, ly, g, H, f0, beta)
    for b in range(args.members):
        pl.set_field(0, h + 1e-3 * normals(3 + b, nx * ny).reshape(ny, nx), b)
        pl.set_field(1, u, b)
    dt = pl.stable_dt(0.5)
    m0 = pl.mean_depth(0)
    pl.steps(dt, args.warmup)
    t0 = time.perf_counter()
    kernel_s = pl.steps_timed(dt, args.steps) / 1e3 / args.steps  # one k_plane_march launch per RK4 step
    wall = time.perf_counter() - t0
    cells = nx * ny
    word = 8 if prec == lmw.FP64 else 4
    bpc = 6.0 * word  # algorithmic bytes per cell per RK4 step: h, u, v read once and written once
    canon = 30.0 * word  # the same step as four stage passes: 4 x (3 fields in + 3 tendencies out) + state in and out
    gbs = bpc * cells * args.members / kernel_s / 1e9
    return {
        "metric": "beta-plane member-days/s (C-grid channel)", "value": args.members * dt * args.steps / 86400.0 / wall,
        "unit": "member-days/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == lmw.FP64 else "f32", "data": "synthetic",
        "config": {"workload": f"beta-plane channel 1024x512 (dx 10 km), sech^2 jet + 1e-3 m noise, batch {args.members}, "
                               f"RK4 dt {dt:.2f} s", "members": args.members, "dt_s": dt},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": gbs * 1e9 / HBM_PEAK,
                     "traffic": measured_traffic(f"plane_{nx}x{ny}x{args.members}_{args.precision}", "k_plane_march"),
                     "kernel": "k_plane_march", "kernel_ms": kernel_s * 1e3,
                     "work_per_launch": f"{bpc * cells * args.members / 1e6:.1f} MB",
                     "per_stage_equivalent_gbs": canon * cells * args.members / kernel_s / 1e9},
        "mass_drift": (pl.mean_depth(0) - m0) / m0,
    }


ENS_GROUPS = ((42, 128, 600.0), (85, 96, 300.0), (170, 32, 150.0))  # truncation, members, dt


def ensemble_state(lmw):
    """The T42 random balanced state (energy up to n = 21, seed 1, no orography): spectral (zeta, delta, phi)."""
    from lmweather_amd.initial import random_balanced_state
    base = lmw.Sphere(42, lmw.SphereParams(A, OM, G))
    h, u, v = random_balanced_state(base, 21, 1)
    base.set_state_grid(h, u, v)
    return [base.get_spectral(f) for f in range(3)]


def build_uneven(lmw, precision, spec, groups=ENS_GROUPS):
    """Members in truncation order; member i: the T42 state embedded at its truncation, vorticity perturbed by 1e-3
    relative noise (seed 1000 + i)."""
    from lmweather_amd.initial import embed, perturb_vorticity
    truncs = [t for t, n, _ in groups for _ in range(n)]
    ens = lmw.Ensemble(truncs, lmw.SphereParams(A, OM, G), precision)
    for i, t in enumerate(truncs):
        for f in range(3):
            ens.set_spectral(i, f, embed(spec[f], t))
        ens.set_spectral(i, 0, perturb_vorticity(ens.get_spectral(i, 0), 1e-3, 1000 + i))
    return ens


def ensemble_bench(lmw, args) -> dict:
    """256 members of uneven truncations (128 x T42 at 600 s, 96 x T85 at 300 s, 32 x T170 at 150 s), stepped in
    600 s macro steps (one HIP graph each: 1, 2 and 4 RK4 steps of the three groups, concurrent on their streams),
    with autotuned Legendre tiles.  value = member-days/s of the whole ensemble."""
    prec = PREC_NAMES.index(args.precision)
    spec = ensemble_state(lmw)
    macro = 600.0
    dts = [dt for _, _, dt in ENS_GROUPS]
    subs = [int(round(macro / dt)) for dt in dts]
    nmembers = sum(n for _, n, _ in ENS_GROUPS)
    days = args.steps * macro / 86400.0

    def run(precision):
        ens = build_uneven(lmw, precision, spec)
        tiles, tune_ms = ens.autotune()
        if args.persistent and precision != lmw.FP64:
            ens.set_persistent(True)
        ens.macro_steps(dts, subs, args.warmup)
        t0 = time.perf_counter()
        ens.macro_steps(dts, subs, args.steps)  # synchronous: returns when the device is done
        wall = time.perf_counter() - t0
        # per-group kernel times of the same kind of step, eagerly with HIP events (groups one after another)
        ms = ens.steps_timed(dts, subs)
        fin = all(np.isfinite(ens.diag(i, lmw._lib.DIAG_ENERGY)) for i in (0, nmembers - 1))
        del ens
        return wall, tiles, tune_ms, ms, fin

    print(f"ensemble {args.precision}: {nmembers} members, {args.steps} macro steps of {macro:g} s", file=sys.stderr, flush=True)
    wall, tiles, tune_ms, ms, fin = run(prec)
    value = nmembers * days / wall
    out = {
        "metric": "ensemble member-days/s, 256 members of uneven truncations (T42/T85/T170)",
        "value": value,
        "unit": "member-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": "synthetic",
        "config": {"workload": "uneven ensemble: 128 x T42 (dt 600 s), 96 x T85 (dt 300 s), 32 x T170 (dt 150 s); "
                               "T42 balanced random flow n<=21 embedded at each truncation, 1e-3 vorticity perturbations; "
                               "one step = one 600 s macro step (HIP graph of 1/2/4 RK4 steps per group)",
                   "members": nmembers, "simulated_days": days, "precision": args.precision},
        "t170_member_days_per_s": 32 * days / wall,
        "legendre_kernels": "persistent multi-group" if args.persistent else "per group, groups as graph branches",
        "finite": bool(fin),
        # FP64: (members per block, 0) of the FMA kernels or (0, row chunk / waves) of the FP64 MFMA kernels
        "tiles": {str(t): ({"fp64_synth_mb_rc": list(lmw.sphere.tile_variant(0, int(tl[0]))),
                            "fp64_anal_mb_nw": list(lmw.sphere.tile_variant(3, int(tl[1])))} if prec == lmw.FP64 else
                           {"mixed_synth_rc": lmw.sphere.tile_variant(1, int(tl[2]))[0],
                            "mixed_anal_kc_nw": list(lmw.sphere.tile_variant(2, int(tl[3])))})
                  for (t, _, _), tl in zip(ENS_GROUPS, tiles)},
        "kernel_ms_per_macro_step": {str(t): {"inverse": float(m[0]), "grid": float(m[1]), "forward": float(m[2])}
                                     for (t, _, _), m in zip(ENS_GROUPS, ms)},
    }
    if prec != lmw.FP64 and not args.only_batch:
        w64, _, _, _, _ = run(lmw.FP64)
        out["fp64_member_days_per_s"] = nmembers * days / w64
        out["speedup_vs_fp64_batched"] = value / out["fp64_member_days_per_s"]
    return out


# sphere_adjustment.lms:18: a Gaussian low in a resting 5960 m layer, plus a zonal wind
FIELDS_EXPR = ("5960.0 - 100.0 * pow(2.718281828459045, -((lon * cos(lat)) * (lon * cos(lat)) + (lat - 0.7853981633974483) * "
               "(lat - 0.7853981633974483)) / 0.04)", "20.0 * cos(lat)", "0")


def setup_bench(lmw, args):
    """The ensemble initial-condition path: set_state_grid of `members` T170 members from host grids, one member per
    call (the reference's shape) against the batched call (lmw_ensemble_set_state_grids), and the device-side
    Williamson 6 set-up (lmw_ensemble_setup: no host grids).  value = member set-ups/s of the batched call; the CPU
    baseline is the reference's own set_state_grid (oracle/_ref, else the oracle) on one core."""
    members = args.members if args.members != 32 else 256
    prec = PREC_NAMES.index(args.precision)
    h, u, v, hs = initial_state(lmw)
    H, U, V = (np.ascontiguousarray(np.broadcast_to(x, (members,) + x.shape)) for x in (h, u, v))
    ens = lmw.Ensemble([N] * members, lmw.SphereParams(A, OM, G), prec)
    ens.set_state_grid(0, h, u, v)  # warm-up: first launches, buffers
    ens.set_state_grids(0, 2, H[:2], U[:2], V[:2])
    ens.setup(0, 2, lmw._lib.SETUP_WILLIAMSON6)
    t0 = time.perf_counter()
    for i in range(members):
        ens.set_state_grid(i, h, u, v)
    t_one = time.perf_counter() - t0
    t0 = time.perf_counter()
    ens.set_state_grids(0, members, H, U, V)
    t_batch = time.perf_counter() - t0
    t0 = time.perf_counter()
    ens.setup(0, members, lmw._lib.SETUP_WILLIAMSON6)
    t_dev = time.perf_counter() - t0
    out = {
        "metric": "ensemble set-up: member set_state_grid/s at T170",
        "value": members / t_batch,
        "unit": "member set-ups/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": 1e3 * t_batch,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{members} T170 members (512x256 Gaussian grid): set_state_grid from host grids of the "
                               "balanced random flow; one step = the whole ensemble set up once", "members": members,
                   "precision": args.precision},
        "one_member_per_call_per_s": members / t_one,
        "device_williamson6_per_s": members / t_dev,
        "batched_speedup": t_one / t_batch,
        "host_grid_bytes": int(H.nbytes * 3),
    }
    del ens
    # the seeded plane perturbation (plane.cpp:34-57) of 64 members of the 1024 x 512 channel, on the device
    pl = lmw.Plane(lmw.PlaneParams(1024, 512, 1.024e7, 5.12e6, 9.81, 1000.0, 1e-4, 1.6e-11), 64, lmw.FP64)
    pl.perturb(1.0, 1, 0)
    t0 = time.perf_counter()
    for b in range(64):
        pl.perturb(1.0, 100 + b, b)
    out["plane_perturb_1024x512_per_s"] = 64 / (time.perf_counter() - t0)
    del pl
    # Sphere.fields (main.cpp:158-172) of every member from expressions: compiled once, evaluated on the device
    ens = lmw.Ensemble([N] * members, lmw.SphereParams(A, OM, G), prec)
    ens.fields(0, 2, *FIELDS_EXPR)
    t0 = time.perf_counter()
    ens.fields(0, members, *FIELDS_EXPR)
    out["device_fields_expr_per_s"] = members / (time.perf_counter() - t0)
    del ens
    if not args.no_cpu:
        from oracle import pyapi
        kind = "reference" if pyapi.available("ref") else "port"
        op = pyapi.OraclePlane(1024, 512, 1.024e7, 5.12e6, 9.81, 1000.0, 1e-4, 1.6e-11, kind="ref" if kind == "reference" else "oracle")
        t0 = time.perf_counter()
        for b in range(3):
            op.perturb(1.0, 100 + b)
        out["cpu_plane_perturb_1024x512_per_s"] = 3 / (time.perf_counter() - t0)
        s = pyapi.OracleSphere(N, A, OM, G, kind="ref" if kind == "reference" else "oracle")
        s.set_state_grid(h, u, v)
        n = 10
        t0 = time.perf_counter()
        for _ in range(n):
            s.set_state_grid(h, u, v)
        wall = time.perf_counter() - t0
        from oracle import expr as OX
        if OX.ref_available():  # the reference's Sphere.fields: its compiler over the grid, then set_state_grid
            lon, lat = OX.sphere_coords(s.gauss()[0], s.nlon)
            LON, LAT = np.meshgrid(lon, lat)
            t0 = time.perf_counter()
            g = [OX.ref_eval(t, ["lon", "lat"], OX.sphere_params(A, OM, G), LON, LAT).reshape(LON.shape) for t in FIELDS_EXPR]
            s.set_state_grid(*g)
            out["cpu_reference_fields_expr_per_s"] = 1 / (time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": n / wall, "unit": "member set-ups/s", "cores": 1, "kind": kind,
                               "sample": f"{n} set_state_grid calls of one T170 member, single thread, {wall:.2f} s"}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="sphere", choices=["sphere", "plane", "ensemble", "setup", "c1", "accuracy", "diagnostics", "corrector"])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--members", type=int, default=32)
    ap.add_argument("--precision", default="bf16x6", choices=list(PREC_NAMES))
    ap.add_argument("--check-days", type=float, default=1.0)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-batch", action="store_true", help="time the batch only (clean kernel profiles)")
    ap.add_argument("--plane-tile", type=int, default=None, help="plane: step kernel variant (Plane.set_tile)")
    ap.add_argument("--persistent", action="store_true", help="ensemble: persistent multi-group Legendre kernels")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding is a design, not a run",
              file=sys.stderr)
        return 2

    import lmweather_amd as lmw
    if args.workload == "plane":
        print(json.dumps(plane_bench(lmw, args)))
        return 0
    if args.workload == "ensemble":
        print(json.dumps(ensemble_bench(lmw, args)))
        return 0
    if args.workload == "setup":
        print(json.dumps(setup_bench(lmw, args)))
        return 0
    if args.workload == "c1":
        print(json.dumps(c1_bench(lmw, args)))
        return 0
    if args.workload == "accuracy":
        print(json.dumps(accuracy_bench(lmw, args)))
        return 0
    if args.workload == "diagnostics":
        print(json.dumps(diagnostics_bench(lmw, args)))
        return 0
    if args.workload == "corrector":
        print(json.dumps(corrector_bench(lmw, args)))
        return 0
    prec = PREC_NAMES.index(args.precision)
    mixed = prec != lmw.FP64
    name, cus, _ = lmw.device_info()
    print(f"device {name} ({cus} CUs); building {args.members} T170 members", file=sys.stderr, flush=True)
    state = initial_state(lmw)
    ens = build_batch(lmw, args.members, prec, state)
    _, J, K = ens.dims(0)
    S = (N + 1) * (N + 2) // 2
    start = [batch_drifts(lmw, ens, args.members)] if mixed and not args.only_batch else None
    wall, ms = timed_batch(ens, args.steps, args.warmup)
    if start is None:
        del ens
    value = args.members * DT * args.steps / 86400.0 / wall
    print(f"{args.precision}: timed {args.steps} steps in {wall:.3f} s", file=sys.stderr, flush=True)

    launches = 4 * args.steps
    synth_f, grid_b, anal_f = kernel_costs(J, K, S, N + 1, args.members, mixed)
    avg = ms / launches * 1e-3  # seconds per launch
    peak = PEAKS[args.precision]
    mult = MFMA_PRODUCTS[args.precision]
    kernels = {
        "legendre_inverse": {"kernel": "k_synth_mx" if mixed else "k_synth", "avg_us": avg[0] * 1e6,
                             "tflops": synth_f / avg[0] / 1e12, "frac": mult * synth_f / avg[0] / peak},
        "grid_fft_products": {"kernel": "k_tend_grid", "avg_us": avg[1] * 1e6, "gbs": grid_b / avg[1] / 1e9,
                              "frac": grid_b / avg[1] / HBM_PEAK},
        "legendre_forward_rk4": {"kernel": "k_anal_mx" if mixed else "k_anal", "avg_us": avg[2] * 1e6,
                                 "tflops": anal_f / avg[2] / 1e12, "frac": mult * anal_f / avg[2] / peak},
    }
    dom = int(np.argmax(ms))
    prof_name = f"sphere_T{N}x{args.members}_{args.precision}"
    if dom == 1:
        roof = {"bound": "hbm", "achieved": grid_b / avg[1] / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": grid_b / avg[1] / HBM_PEAK, "traffic": measured_traffic(prof_name, "k_tend_grid"),
                "kernel": "k_tend_grid", "work_per_launch": f"{grid_b / 1e6:.1f} MB"}
    else:
        fl = synth_f if dom == 0 else anal_f
        kname = kernels[list(kernels)[dom]]["kernel"]
        roof = {"bound": "mfma", "achieved": fl / avg[dom] / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": mult * fl / avg[dom] / peak, "traffic": measured_traffic(prof_name, kname), "kernel": kname,
                "work_per_launch": f"{fl / 1e9:.3f} GFLOP"}

    out = {
        "metric": "ensemble member-days/s at T170, mixed precision vs FP64",
        "value": value,
        "unit": "member-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": "synthetic",
        "config": {"workload": f"T170 ensemble, {args.members} members (512x256 Gaussian grid, dt {DT:g} s, RK4), "
                               "balanced random flow n<=85 over 8 Gaussian mountains, 1e-3 vorticity perturbations",
                   "members": args.members, "truncation": N, "dt_s": DT, "precision": args.precision},
        "roofline": roof,
        "kernels": kernels,
    }
    if not args.only_batch:
        if mixed:
            e64 = build_batch(lmw, args.members, lmw.FP64, state)
            start.append(batch_drifts(lmw, e64, args.members))
            n64 = max(args.steps // 2, 4)
            w64, _ = timed_batch(e64, n64, args.warmup)
            out["fp64_batched_member_days_per_s"] = args.members * DT * n64 / 86400.0 / w64
        p = lmw.SphereParams(A, OM, G)
        one = lmw.Sphere(N, p)
        h, u, v, hs = state
        one.set_orography_grid(hs)
        one.set_state_grid(h, u, v)
        one.steps(DT, 4)
        t1 = time.perf_counter()
        one.steps(DT, args.steps)
        out["fp64_plain_member_days_per_s"] = DT * args.steps / 86400.0 / (time.perf_counter() - t1)
        out["speedup_vs_fp64_plain"] = value / out["fp64_plain_member_days_per_s"]
        if "fp64_batched_member_days_per_s" in out:
            out["speedup_vs_fp64_batched"] = value / out["fp64_batched_member_days_per_s"]
        if mixed:
            print(f"accuracy: {args.check_days} days of the timed batch against the FP64 batch", file=sys.stderr, flush=True)
            from lmweather_amd.initial import perturb_vorticity
            twin = build_batch(lmw, args.members, lmw.FP64, state)
            for i in range(args.members):
                twin.set_spectral(i, 0, perturb_vorticity(twin.get_spectral(i, 0), 1e-7, 77 + i))
            done = (args.warmup + 2 * args.steps, args.warmup + 2 * n64, 0)
            out["accuracy_vs_fp64"] = batch_accuracy(lmw, ens, e64, twin, done, args.members, args.check_days, start)
            del ens, e64, twin
        if not args.no_cpu:
            print(f"cpu baseline: {args.cpu_steps} reference steps", file=sys.stderr, flush=True)
            out["cpu_baseline"] = cpu_baseline(state, args.cpu_steps)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
