"""Seeded initial states of the benchmark configurations (host-side set-up, run once per member).

* ``random_balanced_state``: streamfunction with Gaussian random spherical-harmonic coefficients of amplitude
  proportional to n^-1.5 for 1 <= n <= nmax, scaled so that max |u| on the grid equals ``umax``; height from linear
  balance (laplacian phi' = div(f grad psi)); divergence zero.
* ``gaussian_orography``: a sum of Gaussian bumps hs = H exp(-(r/R)^2), r the great-circle angle to the centre.
* ``perturb_vorticity``: zeta_nm <- zeta_nm (1 + eps xi_nm).

The random numbers are the reference's platform-independent stream (splitmix64 + Irwin-Hall normal,
cpp/sphere.cpp:13-18 and :315-320).  The transforms are those of the ``model`` argument, so the same builder runs on
the GPU Sphere (product path) or on any object with the same transform methods; both paths then receive the
identical grid (h, u, v) through set_state_grid.
"""
from __future__ import annotations

import math

import numpy as np

PHI_GOLDEN = 0x9E3779B97F4A7C15
MASK = (1 << 64) - 1


class SplitMix:
    """splitmix64 with the seeding of the reference's perturb (state = seed * golden + offset)."""

    def __init__(self, seed: int, offset: int = 12345):
        self.x = (seed * PHI_GOLDEN + offset) & MASK

    def next(self) -> int:
        self.x = (self.x + PHI_GOLDEN) & MASK
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return float(self.next() >> 11) * (1.0 / 9007199254740992.0)

    def normal(self) -> float:
        s = 0.0
        for _ in range(12):
            s += self.uniform()
        return s - 6.0


def normals(seed: int, count: int, offset: int = 12345) -> np.ndarray:
    """The first ``count`` Irwin-Hall normals of SplitMix(seed, offset), vectorized (the k-th splitmix state is
    x0 + k * golden mod 2^64, so the whole stream is one array expression); summed left to right like the scalar
    generator, so the values are identical."""
    with np.errstate(over="ignore"):
        x0 = np.uint64((seed * PHI_GOLDEN + offset) & MASK)
        k = np.arange(1, 12 * count + 1, dtype=np.uint64)
        z = x0 + k * np.uint64(PHI_GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u = u.reshape(count, 12)
    s = np.zeros(count)
    for i in range(12):
        s = s + u[:, i]
    return s - 6.0


def random_streamfunction(N: int, nmax: int, seed: int) -> np.ndarray:
    """psi_nm = n^-1.5 (xi_re + i xi_im), 1 <= n <= nmax, xi_im = 0 at m = 0; complex (N+1, N+1) indexed [m, n].
    The normals are drawn in the reference perturb's order (m outer, n inner, re then im)."""
    slots = []
    for m in range(nmax + 1):
        for n in range(max(m, 1), nmax + 1):
            slots.append((m, n, 0))
            if m:
                slots.append((m, n, 1))
    xi = normals(seed, len(slots))
    psi = np.zeros((N + 1, N + 1), dtype=np.complex128)
    for (m, n, part), x in zip(slots, xi):
        if part == 0:
            psi[m, n] += x
        else:
            psi[m, n] += 1j * x
    n = np.arange(N + 1, dtype=np.float64)
    scale = np.zeros(N + 1)
    scale[1:] = n[1:] ** -1.5
    return psi * scale[None, :]


def _lap_factor(N: int, radius: float) -> np.ndarray:
    n = np.arange(N + 1, dtype=np.float64)
    return -n * (n + 1.0) / (radius * radius)


def random_balanced_state(model, nmax: int, seed: int, umax: float = 40.0, depth: float = 8000.0):
    """Grids (h, u, v) of a non-divergent random flow in linear balance (see module docstring).

    ``model`` provides N, radius/omega/gravity (``params``), ``lats()``, ``synthesize_scalar``,
    ``winds_spectral_to_grid`` and ``analyze_divergence``.
    """
    N = model.N
    a, om, g = model.params.radius, model.params.omega, model.params.gravity
    psi = random_streamfunction(N, nmax, seed)
    lap = _lap_factor(N, a)[None, :]
    zeta = psi * lap
    zero = np.zeros_like(zeta)
    lat = model.lats()[:, None]
    cosl = np.cos(lat)
    U, V = model.winds_spectral_to_grid(zeta, zero)
    speed = np.hypot(U, V) / cosl
    scale = umax / speed.max()
    psi, zeta = psi * scale, zeta * scale
    U, V = U * scale, V * scale
    # linear balance: laplacian(phi') = div(f grad psi); cos(lat) grad psi = the irrotational wind of potential psi
    Ug, Vg = model.winds_spectral_to_grid(zero, zeta)
    f = 2 * om * np.sin(lat)
    div = model.analyze_divergence(f * Ug, f * Vg)
    inv = np.zeros(N + 1)
    inv[1:] = 1.0 / _lap_factor(N, a)[1:]
    phi = div * inv[None, :]
    phi[0, 0] = g * depth * math.sqrt(2.0)
    h = model.synthesize_scalar(phi) / g
    return h, U / cosl, V / cosl


def gaussian_orography(model, count: int = 8, seed: int = 2, hmin: float = 500.0, hmax: float = 2000.0,
                       rmin_deg: float = 10.0, rmax_deg: float = 20.0) -> np.ndarray:
    """Sum of ``count`` Gaussian bumps with heights U[hmin, hmax] m, radii U[rmin, rmax] degrees, centres uniform on
    the sphere; grid of shape (nlat, nlon)."""
    rng = SplitMix(seed, offset=777)
    lat = model.lats()[:, None]
    lon = model.lons()[None, :]
    hs = np.zeros((lat.size, lon.size))
    for _ in range(count):
        hb = hmin + (hmax - hmin) * rng.uniform()
        rad = math.radians(rmin_deg + (rmax_deg - rmin_deg) * rng.uniform())
        lonc = 2 * math.pi * rng.uniform()
        latc = math.asin(2 * rng.uniform() - 1)
        c = np.sin(lat) * math.sin(latc) + np.cos(lat) * math.cos(latc) * np.cos(lon - lonc)
        r = np.arccos(np.clip(c, -1.0, 1.0))
        hs += hb * np.exp(-(r / rad) ** 2)
    return hs


def perturb_vorticity(zeta: np.ndarray, rel: float, seed: int) -> np.ndarray:
    """zeta_nm (1 + rel xi_nm) with xi ~ N(0, 1) from the seeded stream, over the stored triangle n >= m (m outer)."""
    N = zeta.shape[0] - 1
    m_idx, n_idx = np.triu_indices(N + 1)
    xi = normals(seed, m_idx.size)
    out = zeta.copy()
    out[m_idx, n_idx] = zeta[m_idx, n_idx] * (1.0 + rel * xi)
    return out


def embed(spec: np.ndarray, N: int) -> np.ndarray:
    """Spectral coefficients copied into truncation N (higher n dropped or zero-filled): the normalization of the
    orthonormal basis does not depend on the truncation."""
    out = np.zeros((N + 1, N + 1), dtype=np.complex128)
    k = min(N, spec.shape[0] - 1)
    out[: k + 1, : k + 1] = spec[: k + 1, : k + 1]
    return out
