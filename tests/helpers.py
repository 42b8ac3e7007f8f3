"""Helpers shared by the tests: adapters that let the initial-condition builders drive the oracle, and norms."""
import numpy as np


def rel_l2(a, b) -> float:
    a, b = np.asarray(a), np.asarray(b)
    den = np.linalg.norm(np.ravel(b))
    return float(np.linalg.norm(np.ravel(a - b)) / (den if den > 0 else 1.0))


def max_rel(a, b, scale=None) -> float:
    """max |a - b| relative to max |b| (or to ``scale`` when b is only round-off, e.g. the divergence of a
    non-divergent flow)."""
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(b)) if scale is None else scale
    return float(np.max(np.abs(a - b)) / max(s, 1e-300))


class OracleTransforms:
    """Presents an oracle.pyapi.OracleSphere with the transform method names of lmweather_amd.Sphere."""

    def __init__(self, s, params):
        self.s, self.params, self.N = s, params, s.N

    def lats(self):
        return np.arcsin(self.s.gauss()[0])

    def lons(self):
        return 2 * np.pi * np.arange(self.s.nlon) / self.s.nlon

    def synthesize_scalar(self, spec):
        return self.s.synthesize(spec)

    def winds_spectral_to_grid(self, z, d):
        return self.s.winds(z, d)

    def analyze_divergence(self, A, B):
        return self.s.analyze_divergence(A, B)


def random_spec(N, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    s = np.zeros((N + 1, N + 1), dtype=np.complex128)
    for m in range(N + 1):
        for n in range(m, N + 1):
            s[m, n] = scale * (rng.standard_normal() + (1j * rng.standard_normal() if m else 0.0))
    return s
