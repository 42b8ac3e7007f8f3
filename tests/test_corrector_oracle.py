"""Pins the CPU restatement of the corrector network (oracle/corrector_net.py) before it is used to check the GPU:
hand-derived gradients against central finite differences, BF16 rounding against known answers and torch's
conversion, pooling against brute force, Adam against torch.optim.Adam."""
import numpy as np
import pytest

from oracle import corrector_net as cn


def _problem(N=6, W=8, S=2, seed=0):
    rng = np.random.default_rng(seed)
    _, n = cn.layout(W)
    params = rng.standard_normal(n) * 0.4
    C = cn.tri_size(N)
    x = rng.standard_normal((S, 3, C, 2))
    t = rng.standard_normal((S, 3, C, 2)) * 0.5
    ins = rng.uniform(0.5, 2.0, 3 * (N + 1))
    outs = rng.uniform(0.5, 2.0, 3 * (N + 1))
    return params, x, t, ins, outs


def test_gradient_matches_finite_differences():
    N, W = 6, 8
    params, x, t, ins, outs = _problem(N, W)
    loss, g = cn.loss_grad(params, W, N, ins, outs, x, t, round_weights=False)
    assert loss > 0
    rng = np.random.default_rng(1)
    idx = rng.choice(params.size, 60, replace=False)
    lay, _ = cn.layout(W)
    # every tensor is probed at least once
    idx = np.unique(np.concatenate([idx, [o for o, _ in lay.values()]]))
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (cn.loss_grad(p1, W, N, ins, outs, x, t, False)[0] - cn.loss_grad(p2, W, N, ins, outs, x, t, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_gradient_wrong_sign_or_dropped_term_is_caught():
    """A gradient with the mixing path removed differs from finite differences: the check above has teeth."""
    N, W = 6, 8
    params, x, t, ins, outs = _problem(N, W)
    lay, _ = cn.layout(W)
    _, g = cn.loss_grad(params, W, N, ins, outs, x, t, round_weights=False)
    o, shp = lay["W2"]
    p2 = params.copy()
    p2[lay["V2"][0]:lay["V2"][0] + W * W] = 0.0  # no mixing: the W2 gradient must change
    _, g2 = cn.loss_grad(p2, W, N, ins, outs, x, t, round_weights=False)
    assert np.abs(g2[o:o + W * W] - g[o:o + W * W]).max() > 1e-6


def test_bf16_rounding_known_answers():
    one = 1.0
    cases = {one: one, one + 2.0**-8: one, one + 3 * 2.0**-8: one + 2.0**-6, one + 2.0**-7: one + 2.0**-7,
             -(one + 2.0**-8 + 2.0**-20): -(one + 2.0**-7), 2.0**-130 + 2.0**-140: 2.0**-130}
    for a, want in cases.items():
        got = cn.bf16_round(np.float32(a))
        assert got == np.float64(np.float32(want)), (a, got, want)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    v = np.random.default_rng(3).standard_normal(100000).astype(np.float32) * np.float32(10.0) ** np.random.default_rng(4).integers(-20, 20, 100000)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(cn.bf16_round(v), ref)


def test_pooling_brute_force():
    N, S, F = 5, 2, 3
    C = cn.tri_size(N)
    H = np.random.default_rng(5).standard_normal((S * C, F))
    P = cn._pool(H, S, N)
    deg, ordm = cn.tri_tables(N)
    for s in range(S):
        for n in range(N + 1):
            rows = [s * C + c for c in range(C) if deg[c] == n]
            assert len(rows) == n + 1
            assert np.allclose(P[s, n], H[rows].mean(0), rtol=0, atol=1e-14)
    # packed order: index m (N+1) - m (m-1)/2 + n - m
    for c in range(C):
        m, n = ordm[c], deg[c]
        assert c == m * (N + 1) - m * (m - 1) // 2 + n - m


def test_mask_and_correction_zero_where_untrained():
    N, W = 6, 8
    params, x, _, ins, outs = _problem(N, W)
    corr, res = cn.correction(params, W, N, ins, outs, x)
    assert np.all(corr[:, :, 0, :] == 0)  # (0, 0) row
    _, ordm = cn.tri_tables(N)
    assert np.all(corr[:, :, ordm == 0, 1] == 0)  # imaginary parts at m = 0
    assert res.shape == (2,) and np.all(res > 0)
    assert cn.trained_mask(N).sum() == (cn.tri_size(N) - 1) * 6 - 3 * N


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(6)
    p0 = rng.standard_normal(50)
    grads = [rng.standard_normal(50) for _ in range(5)]
    tp = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m1, m2 = p0.copy(), np.zeros(50), np.zeros(50)
    for i, g in enumerate(grads):
        opt.zero_grad()
        tp.grad = torch.tensor(g)
        opt.step()
        p, m1, m2 = cn.adam(p, g, m1, m2, i + 1, 1e-2)
    assert np.allclose(p, tp.detach().numpy(), rtol=0, atol=1e-12)
