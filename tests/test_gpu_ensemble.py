"""GPU ensembles of uneven truncations: the graph-captured macro step against per-group stepping, every tile variant
of the Legendre kernels against the default shapes, and the autotuner (state untouched, cache reused)."""
import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

EARTH = (6.371e6, 7.292e-5, 9.80616)


@pytest.fixture(scope="module")
def lmw():
    import lmweather_amd
    return lmweather_amd


def build(lmw, truncs, precision):
    from lmweather_amd.initial import embed, perturb_vorticity, random_balanced_state
    e = lmw.Ensemble(truncs, lmw.SphereParams(*EARTH), precision)
    base = lmw.Sphere(42, lmw.SphereParams(*EARTH))
    h, u, v = random_balanced_state(base, 21, 1)
    base.set_state_grid(h, u, v)
    spec = [base.get_spectral(f) for f in range(3)]
    for i, N in enumerate(truncs):
        for f in range(3):
            e.set_spectral(i, f, embed(spec[f], N))
        e.set_spectral(i, 0, perturb_vorticity(e.get_spectral(i, 0), 1e-3, 1000 + i))
    return e


def states(e):
    return [np.stack([e.get_spectral(i, f) for f in range(3)]) for i in range(len(e.truncations))]


@pytest.mark.parametrize("precision,count", [(0, 3), (2, 3), (0, 9), (3, 9)])
def test_macro_steps_equal_group_steps(lmw, precision, count):
    """count 9: two replays of the graph of 4 macro steps, then one of the single macro step."""
    truncs = [42] * 5 + [85] * 3 + [170] * 2
    a, b = build(lmw, truncs, precision), build(lmw, truncs, precision)
    a.macro_steps([600.0, 300.0, 150.0], [1, 2, 4], count)
    b.steps([600.0, 300.0, 150.0], [count, 2 * count, 4 * count])
    for x, y in zip(states(a), states(b)):
        assert np.array_equal(x, y)
    assert a.diag(9, lmw._lib.DIAG_TIME) == 600.0 * count and a.diag(0, lmw._lib.DIAG_TIME) == 600.0 * count


@pytest.mark.parametrize("precision", [2, 3])
def test_persistent_macro_steps_equal_branch_macro_steps(lmw, precision):
    """The persistent multi-group Legendre kernels do per member exactly what the per-group kernels do."""
    truncs = [42] * 5 + [85] * 3 + [170] * 2
    a, b = build(lmw, truncs, precision), build(lmw, truncs, precision)
    a.set_persistent(True)
    a.macro_steps([600.0, 300.0, 150.0], [1, 2, 4], 3)
    b.macro_steps([600.0, 300.0, 150.0], [1, 2, 4], 3)
    for x, y in zip(states(a), states(b)):
        assert np.array_equal(x, y)
    assert a.diag(9, lmw._lib.DIAG_TIME) == 1800.0
    with pytest.raises(ValueError, match="divide the largest"):
        a.macro_steps([600.0, 200.0, 150.0], [1, 3, 4], 1)


def test_persistent_many_tiles_equal_branch_macro_steps(lmw):
    """Groups of several 16-member tiles, the last one partial (20 x T42, 17 x T85, 18 x T170): the blocks of the
    persistent kernels walk lists of several items, the T170 inverse transform's items cover half the latitudes, and
    the branch mode steps each group as concurrent member parts; every member bit-identical."""
    truncs = [42] * 20 + [85] * 17 + [170] * 18
    a, b = build(lmw, truncs, 3), build(lmw, truncs, 3)
    a.set_persistent(True)
    a.macro_steps([600.0, 300.0, 150.0], [1, 2, 4], 2)
    b.macro_steps([600.0, 300.0, 150.0], [1, 2, 4], 2)
    for x, y in zip(states(a), states(b)):
        assert np.array_equal(x, y)


def test_every_tile_variant_matches_default(lmw):
    """Mixed BF16x3: each (row chunk) x (latitude chunk, waves) variant stays within the BF16x3 error of the default
    shapes after 4 steps; FP64: the members-per-block variants are bit-identical (members are independent)."""
    from lmweather_amd.sphere import tile_variant
    truncs = [85] * 20
    ref = build(lmw, truncs, 2)
    ref.steps([300.0], [4])
    r = states(ref)
    seen = set()
    for si in range(2):
        for ai in range(6):
            e = build(lmw, truncs, 2)
            e.set_tiles(0, [1, 1, si, ai])
            e.steps([300.0], [4])
            err = max(rel_l2(x[2], y[2]) for x, y in zip(states(e), r))
            print(f"synth RC {tile_variant(1, si)[0]} anal KC/NW {tile_variant(2, ai)}: rel-L2(phi) vs default {err:.2e}")
            assert err <= 1e-6
            seen.add((si, ai))
    assert len(seen) == 12
    f_ref = build(lmw, truncs, 0)
    f_ref.steps([300.0], [2])
    for v in range(3):
        e = build(lmw, truncs, 0)
        e.set_tiles(0, [v, v, 0, 1])
        e.steps([300.0], [2])
        for x, y in zip(states(e), states(f_ref)):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("N", [85, 42])
def test_fp64_mfma_transforms(lmw, N):
    """The FP64 transforms on the FP64 MFMA (FP64 synthesis variants 3, 4: row chunks 32, 64; analysis variants 3, 4:
    4 and 8 waves) against the FMA kernels: the same FP64 sums in another order, so the states agree to rounding after
    4 steps, and so do the state analyses of set_state_grid (the forward transform's STATE mode) and the diagnostic
    grids (the inverse transform outside the step).  9 members: a partial 8-member tile of the synthesis."""
    from lmweather_amd.initial import random_balanced_state
    from lmweather_amd.sphere import tile_variant
    assert tile_variant(3, 1) == (4, 0) and tile_variant(3, 3) == (0, 4) and tile_variant(3, 4) == (0, 8)
    assert tile_variant(0, 1) == (4, 0) and tile_variant(0, 3) == (0, 32) and tile_variant(0, 4) == (0, 64)
    truncs = [N] * 9
    ref = build(lmw, truncs, 0)
    ref.steps([300.0], [4])
    base = lmw.Sphere(N, lmw.SphereParams(*EARTH))
    h, u, v = random_balanced_state(base, N // 2, 7)
    g_ref = build(lmw, truncs, 0)
    g_ref.set_state_grid(2, h, u, v)
    for sv, av in ((1, 3), (1, 4), (3, 1), (4, 1), (4, 3)):
        e = build(lmw, truncs, 0)
        e.set_tiles(0, [sv, av, 0, 1])
        e.steps([300.0], [4])
        for x, y in zip(states(e), states(ref)):
            assert rel_l2(x, y) <= 1e-13, (sv, av)
        for f in range(3):
            assert rel_l2(e.grid(8, f), ref.grid(8, f)) <= 1e-13, (sv, av, f)
        g = build(lmw, truncs, 0)
        g.set_tiles(0, [sv, av, 0, 1])
        g.set_state_grid(2, h, u, v)
        zeta = np.linalg.norm(g_ref.get_spectral(2, 0))
        for f in range(3):
            a, b = g.get_spectral(2, f), g_ref.get_spectral(2, f)
            # delta of the balanced state is rounding noise (1e-17 of zeta): measured against zeta's size
            scale = np.linalg.norm(b) if f != 1 else zeta
            assert np.linalg.norm(a - b) <= 1e-13 * scale, (sv, av, f)


def test_autotune_keeps_state_and_uses_cache(lmw, tmp_path, monkeypatch):
    monkeypatch.setenv("LMW_TUNE_CACHE", str(tmp_path / "tiles.txt"))
    truncs = [42] * 16 + [170] * 4
    e = build(lmw, truncs, 2)
    before = states(e)
    tiles, ms = e.autotune(force=True)
    print("tiles", tiles.tolist(), "ms", ms.tolist())
    assert tiles.shape == (2, 4) and np.all(ms > 0)
    for x, y in zip(states(e), before):
        assert np.array_equal(x, y)
    lines = (tmp_path / "tiles.txt").read_text().split("\n")
    assert sum(1 for ln in lines if ln.strip()) == 2
    e2 = build(lmw, truncs, 2)
    tiles2, ms2 = e2.autotune()
    assert np.array_equal(tiles2, tiles) and np.all(ms2 == 0)
    # tuned shapes step to the same answers as the defaults, within BF16x3 error
    d = build(lmw, truncs, 2)
    e2.steps([600.0, 150.0], [2, 8])
    d.steps([600.0, 150.0], [2, 8])
    for x, y in zip(states(e2), states(d)):
        assert rel_l2(x[2], y[2]) <= 1e-6


@pytest.mark.parametrize("precision,n,tiles", [(0, 32, [3, 3, 0, 0]), (0, 20, [1, 1, 0, 0]), (3, 20, [1, 1, 0, 0])])
def test_untuned_default_tiles(lmw, precision, n, tiles):
    """Untuned groups: FP64 groups of 32 members step with the FP64 MFMA kernels (row chunk 32, 4 waves), smaller
    FP64 groups with the FMA kernels of 4 members per block, mixed groups with row chunk 32 and latitude chunk 32
    on 4 waves: the default stepping is bit-identical to those variants set explicitly."""
    truncs = [42] * n
    a, b = build(lmw, truncs, precision), build(lmw, truncs, precision)
    b.set_tiles(0, tiles)
    a.steps([600.0], [3])
    b.steps([600.0], [3])
    for x, y in zip(states(a), states(b)):
        assert np.array_equal(x, y)
