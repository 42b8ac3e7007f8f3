"""Benchmark: ensemble member-days/s at T170 on one MI355X.

Workload (one "step" = one RK4 step of the whole batch): ``--members`` members at T170 (512 x 256 Gaussian grid,
a = 6.371e6 m, dt = 150 s) -- the T170 share of the 256-member uneven ensemble -- each the T170 random balanced state
(energy up to n = 85, seed 1) over 8 seeded Gaussian mountains, with its vorticity coefficients perturbed by 1e-3
relative noise (seed 1000 + member).  Inputs are resident in HBM before the timed region.

Prints one JSON line: value = whole-job member-days/s, the roofline of the dominant kernel measured with HIP events
around every launch on the library's stream, the FP64 single-member ("plain") rate, and the reference CPU code timed
on this host (oracle/_ref, one core) on a bounded sample of the same workload.
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
PEAKS = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16x3": 2.5e15}
HBM_PEAK = 8.0e12


def build_state(lmw, members, precision):
    from lmweather_amd.initial import gaussian_orography, perturb_vorticity, random_balanced_state
    p = lmw.SphereParams(A, OM, G)
    base = lmw.Sphere(N, p)
    h, u, v = random_balanced_state(base, 85, 1)
    hs = gaussian_orography(base, 8, 2)
    ens = lmw.Ensemble([N] * members, p, precision)
    for i in range(members):
        ens.set_orography_grid(i, hs)
        ens.set_state_grid(i, h - hs, u, v)
        ens.set_spectral(i, 0, perturb_vorticity(ens.get_spectral(i, 0), 1e-3, 1000 + i))
    return ens, (h - hs, u, v, hs)


def kernel_costs(J, K, S, M1, members):
    """Algorithmic work per launch (one RK4 stage of the batch).  Legendre flops counted as executed with the
    equatorial folding (half of the reference's 24 J S and 28 J S); the grid kernel's HBM bytes: Fourier
    coefficients in (4 fields), f and hs grids in, analysis operands out."""
    Jh = J // 2
    synth_flops = 24.0 * Jh * S * members
    anal_flops = 28.0 * Jh * S * members
    grid_bytes = (4 * J * M1 * 16 + 2 * J * K * 8 + M1 * Jh * 16 * 16) * members
    return synth_flops, grid_bytes, anal_flops


def cpu_baseline(state, steps):
    """The reference's own code (oracle/_ref, else the oracle restatement) on one host core, `steps` RK4 steps of
    one T170 member from the same initial state."""
    from oracle import pyapi
    kind = "reference" if pyapi.available("ref") else "port"
    h, u, v, hs = state
    s = pyapi.OracleSphere(N, A, OM, G, kind="ref" if kind == "reference" else "oracle")
    s.set_orography_grid(hs)
    s.set_state_grid(h, u, v)
    s.step(DT)  # warm caches
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(DT)
    wall = time.perf_counter() - t0
    return {"value": DT * steps / 86400.0 / wall, "unit": "member-days/s", "cores": 1, "kind": kind,
            "sample": f"{steps} RK4 steps (dt {DT:g} s) of one T170 member, single thread, {wall:.2f} s",
            "ms_per_step": 1e3 * wall / steps}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--members", type=int, default=32)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16x3"])
    ap.add_argument("--cpu-steps", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the single-member FP64 rate (clean kernel profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding is a design, not a run",
              file=sys.stderr)
        return 2

    import lmweather_amd as lmw
    prec = {"fp64": lmw.FP64, "fp32": lmw.FP32, "bf16x3": lmw.BF16X3}[args.precision]
    name, cus, _ = lmw.device_info()
    print(f"device {name} ({cus} CUs); building {args.members} T170 members", file=sys.stderr, flush=True)
    ens, state = build_state(lmw, args.members, prec)
    _, J, K = ens.dims(0)
    S = (N + 1) * (N + 2) // 2
    ens.steps([DT], [args.warmup])
    print(f"warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    ms = ens.steps_timed([DT], [args.steps])[0]  # synchronous: returns after the device finished
    wall = time.perf_counter() - t0
    value = args.members * DT * args.steps / 86400.0 / wall
    print(f"timed {args.steps} steps in {wall:.3f} s", file=sys.stderr, flush=True)

    launches = 4 * args.steps
    synth_f, grid_b, anal_f = kernel_costs(J, K, S, N + 1, args.members)
    avg = ms / launches * 1e-3  # seconds per launch
    peak = PEAKS[args.precision]
    kernels = {
        "legendre_inverse": {"avg_us": avg[0] * 1e6, "tflops": synth_f / avg[0] / 1e12, "frac": synth_f / avg[0] / peak},
        "grid_fft_products": {"avg_us": avg[1] * 1e6, "gbs": grid_b / avg[1] / 1e9, "frac": grid_b / avg[1] / HBM_PEAK},
        "legendre_forward_rk4": {"avg_us": avg[2] * 1e6, "tflops": anal_f / avg[2] / 1e12, "frac": anal_f / avg[2] / peak},
    }
    dom = int(np.argmax(ms))
    if dom == 1:
        roof = {"bound": "hbm", "achieved": grid_b / avg[1] / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": grid_b / avg[1] / HBM_PEAK, "traffic": None, "kernel": "k_tend_grid"}
    else:
        fl = synth_f if dom == 0 else anal_f
        roof = {"bound": "mfma", "achieved": fl / avg[dom] / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": fl / avg[dom] / peak, "traffic": None, "kernel": "k_synth" if dom == 0 else "k_anal"}

    plain = None
    if not args.no_plain:  # FP64 single member ("plain" form of config C2)
        p = lmw.SphereParams(A, OM, G)
        one = lmw.Sphere(N, p)
        h, u, v, hs = state
        one.set_orography_grid(hs)
        one.set_state_grid(h, u, v)
        one.steps(DT, 4)
        t1 = time.perf_counter()
        one.steps(DT, args.steps)
        plain = DT * args.steps / 86400.0 / (time.perf_counter() - t1)

    out = {
        "metric": "ensemble member-days/s at T170",
        "value": value,
        "unit": "member-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 Legendre / f64", "bf16x3": "bf16x3 Legendre / f64"}[args.precision],
        "data": "synthetic",
        "config": {"workload": f"T170 ensemble, {args.members} members (512x256 Gaussian grid, dt {DT:g} s, RK4), "
                               "balanced random flow n<=85 over 8 Gaussian mountains, 1e-3 vorticity perturbations",
                   "members": args.members, "truncation": N, "dt_s": DT, "precision": args.precision},
        "roofline": roof,
        "kernels": kernels,
        "fp64_plain_member_days_per_s": plain,
        # the reference's Legendre flop count (52 J S per member and stage, twice the folded work executed here)
        "reference_equivalent_legendre_tflops": 2 * (synth_f + anal_f) * launches / ((ms[0] + ms[2]) * 1e-3) / 1e12,
    }
    if not args.no_cpu:
        print(f"cpu baseline: {args.cpu_steps} reference steps", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(state, args.cpu_steps)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
