"""Writes the golden vectors of tests/golden/ from the compiled reference (oracle/_ref/liblmw_ref.so, built by
oracle/Makefile from the reference's own cpp/sphere.cpp and cpp/plane.cpp).

Run in a container where the reference tree exists:  python tests/golden/make_golden.py
The outputs are data only (inputs and the reference's outputs); tests/test_oracle_golden.py checks the oracle
against them, so the oracle stays pinned where the reference cannot be built (e.g. on the GPU host).
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyapi  # noqa: E402
from tests.helpers import OracleTransforms, random_spec  # noqa: E402
from lmweather_amd._lib import SphereParams  # noqa: E402
from lmweather_amd.initial import random_balanced_state, gaussian_orography  # noqa: E402

A, OM, G = 6.371e6, 7.292e-5, 9.80616


def main() -> None:
    pyapi.build(ref=True)
    assert pyapi.available("ref"), "oracle/_ref/liblmw_ref.so is needed"
    out: dict[str, np.ndarray] = {}
    scalars: dict[str, object] = {}

    # free functions (sphere.cpp:21-73)
    n = 64
    k = np.arange(n)
    x = np.sin(0.3 * k) + 0.1 * k + 1j * np.cos(1.7 * k)  # the input of tests/core.cpp:23
    out["fft_in"] = x
    out["fft_fwd"] = pyapi.fft(x, -1, "ref")
    out["fft_bwd"] = pyapi.fft(out["fft_fwd"], 1, "ref")
    for nn in (32, 64, 256):
        out[f"gauss_x_{nn}"], out[f"gauss_w_{nn}"] = pyapi.gauss_legendre(nn, "ref")
    out["legcol_m5_mu03"] = pyapi.legendre_column(5, 42, 0.3, "ref")
    out["legcol_m160_mu099"] = pyapi.legendre_column(160, 170, 0.999, "ref")

    # sphere T21: tables, transforms, perturb stream
    s = pyapi.OracleSphere(21, A, OM, G, kind="ref")
    P, H = s.tables()
    out["t21_P"], out["t21_H"] = P, H
    spec = random_spec(21, 3)
    out["t21_spec"] = spec
    out["t21_synth"] = s.synthesize(spec)
    out["t21_analyze"] = s.analyze(out["t21_synth"])
    rng = np.random.default_rng(4)
    Ag, Bg = rng.standard_normal((s.nlat, s.nlon)), rng.standard_normal((s.nlat, s.nlon))
    out["t21_divA"], out["t21_divB"] = Ag, Bg
    out["t21_div"] = s.analyze_divergence(Ag, Bg)
    hrest = np.full((s.nlat, s.nlon), 5000.0)
    s.set_state_grid(hrest, 0 * hrest, 0 * hrest)
    s.perturb(20.0, 3, 5)
    out["t21_perturb_phi"] = s.get_spectral(2)

    # sphere T42 config C1 (a = 6.371e6, random balanced flow, seed 1, n <= 21): tendencies, 1 and 144 steps
    p = SphereParams(A, OM, G)
    s = pyapi.OracleSphere(42, A, OM, G, kind="ref")
    h, u, v = random_balanced_state(OracleTransforms(s, p), 21, 1)
    s.set_state_grid(h, u, v)
    dz, dd, df = s.tendencies()
    out["c1_tend_z"], out["c1_tend_d"], out["c1_tend_f"] = dz, dd, df
    scalars["c1_stable_dt"] = s.stable_dt()
    scalars["c1_energy0"] = s.energy()
    scalars["c1_enstrophy0"] = s.potential_enstrophy()
    s.step(600.0)
    out["c1_step1_z"], out["c1_step1_d"], out["c1_step1_f"] = (s.get_spectral(f) for f in range(3))
    for _ in range(143):
        s.step(600.0)
    out["c1_step144_z"], out["c1_step144_d"], out["c1_step144_f"] = (s.get_spectral(f) for f in range(3))
    scalars["c1_step144_energy"] = s.energy()
    scalars["c1_step144_enstrophy"] = s.potential_enstrophy()
    scalars["c1_step144_max_wind"] = s.max_wind()
    l2, linf = s.height_errors()
    scalars["c1_step144_err_l2"], scalars["c1_step144_err_linf"] = l2, linf

    # sphere T21 with orography, tilt and diffusion, 10 steps
    s = pyapi.OracleSphere(21, A, OM, G, 1e16, kind="ref")
    h, u, v = random_balanced_state(OracleTransforms(s, p), 10, 5)
    hs = gaussian_orography(OracleTransforms(s, p), 8, 2)
    s.set_orography_grid(hs)
    s.set_axis_tilt(0.4)
    s.set_state_grid(h - hs, u, v)
    for _ in range(10):
        s.step(1200.0)
    out["t21_oro_z"], out["t21_oro_d"], out["t21_oro_f"] = (s.get_spectral(f) for f in range(3))
    scalars["t21_oro_potential_energy"] = s.potential_energy()

    # run() messages (sphere.cpp:463-469)
    try:
        s.run(3600.0, 1e5)
    except ValueError as e:
        scalars["sphere_run_refusal"] = str(e)

    # plane 64 x 32: perturb stream, tendency, 1 and 100 steps (plane.cpp:34-158)
    pl = pyapi.OraclePlane(64, 32, 6.4e5, 3.2e5, 9.81, 1000.0, 1e-4, 1.6e-11, 20.0, 1e-7, False, kind="ref")
    pl.perturb(0.5, 11)
    out["plane_perturb_h"] = pl.get_field(0)
    rng = np.random.default_rng(12)
    pu, pv = 0.5 * rng.standard_normal((32, 64)), 0.5 * rng.standard_normal((33, 64))
    pl.set_field(1, pu)
    pl.set_field(2, pv)
    out["plane_u0"], out["plane_v0"] = pl.get_field(1), pl.get_field(2)
    th, tu, tv = pl.tendency()
    out["plane_tend_h"], out["plane_tend_u"], out["plane_tend_v"] = th, tu, tv
    pl.step(30.0)
    out["plane_step1_h"], out["plane_step1_u"], out["plane_step1_v"] = (pl.get_field(f) for f in range(3))
    for _ in range(99):
        pl.step(30.0)
    out["plane_step100_h"], out["plane_step100_u"], out["plane_step100_v"] = (pl.get_field(f) for f in range(3))
    scalars["plane_step100_energy"] = pl.energy()
    scalars["plane_step100_pv_rms"] = pl.pv_anomaly_rms()
    scalars["plane_stable_dt"] = pl.stable_dt()
    try:
        pl.run(1e5, 1e4)
    except ValueError as e:
        scalars["plane_run_refusal"] = str(e)

    npz = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(npz, **out)
    with open(os.path.join(HERE, "reference_scalars.json"), "w") as f:
        json.dump(scalars, f, indent=1, sort_keys=True)
    manifest = {}
    for name in ("reference_vectors.npz", "reference_scalars.json"):
        with open(os.path.join(HERE, name), "rb") as f:
            manifest[name] = hashlib.sha256(f.read()).hexdigest()
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print({k: v.shape for k, v in out.items()})
    print(scalars)


if __name__ == "__main__":
    main()
