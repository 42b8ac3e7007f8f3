"""CPU restatement (numpy, float64) of the learned coarse-step corrector's network: forward pass, loss, backward
pass and Adam.  TEST INFRASTRUCTURE ONLY: the tests import it to check the GPU kernels' outputs and gradients; the
product path never imports it.

The network has no counterpart in the reference: it is an addition beside the reference's model.  Its
definition, as implemented by the GPU kernels:

* rows are (sample, packed coefficient (m, n)), C = (N+1)(N+2)/2 per sample, packed index
  m (N+1) - m (m-1)/2 + n - m;
* features X [rows][8] = (re, im) of zeta, delta, phi divided by in_scale[field][n], then n/N, m/N;
* H1 = tanh(X W1^T + b1); H2 = tanh(H1 W2^T + b2 + mean_m(H1)[n] V2^T); H3 likewise with W3, b3, V3 on H2;
  Y = H3 W4^T + b4 (6 outputs); products use the weights rounded to BF16 (round to nearest even), the biases as
  they are;
* loss = mean over trained entries of (Y - target / out_scale)^2; trained entries exclude the (0, 0) row and the
  imaginary parts at m = 0.

The gradient here is derived by hand like the GPU's, and the tests pin it against central finite differences in
float64 (with rounding off, so the loss is smooth in the weights).
"""
from __future__ import annotations

import numpy as np

FIN, FOUT = 8, 6


def tri_size(N: int) -> int:
    return (N + 1) * (N + 2) // 2


def tri_tables(N: int):
    """(deg, ordm) of each packed index."""
    deg = np.empty(tri_size(N), dtype=np.int64)
    ordm = np.empty_like(deg)
    c = 0
    for m in range(N + 1):
        for n in range(m, N + 1):
            deg[c], ordm[c] = n, m
            c += 1
    return deg, ordm


def bf16_round(a) -> np.ndarray:
    """float32 -> bfloat16 with round to nearest, ties to even (finite inputs), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout(W: int):
    """Offsets of the flat parameter vector: W1 [W][8], b1, W2 [W][W], b2, V2, W3, b3, V3, W4 [6][W], b4."""
    shapes = [("W1", (W, FIN)), ("b1", (W,)), ("W2", (W, W)), ("b2", (W,)), ("V2", (W, W)), ("W3", (W, W)), ("b3", (W,)),
              ("V3", (W, W)), ("W4", (FOUT, W)), ("b4", (FOUT,))]
    out, o = {}, 0
    for name, shp in shapes:
        n = int(np.prod(shp))
        out[name] = (o, shp)
        o += n
    return out, o


def unpack(params, W: int, round_weights: bool = True):
    lay, n = layout(W)
    params = np.asarray(params)
    if params.size != n:
        raise ValueError("parameter count mismatch")
    p = {}
    for name, (o, shp) in lay.items():
        v = params[o:o + int(np.prod(shp))].reshape(shp)
        p[name] = bf16_round(v) if (round_weights and name[0] in "WV") else v.astype(np.float64)
    return p


def pack(grads: dict, W: int) -> np.ndarray:
    lay, n = layout(W)
    out = np.zeros(n)
    for name, (o, shp) in lay.items():
        out[o:o + int(np.prod(shp))] = grads[name].ravel()
    return out


def trained_mask(N: int) -> np.ndarray:
    """(C, 6) bool: entries the loss trains and the correction touches."""
    _, ordm = tri_tables(N)
    m = np.ones((tri_size(N), FOUT), dtype=bool)
    m[0, :] = False
    m[ordm == 0, 1::2] = False
    return m


def features(raw, in_scale, N: int) -> np.ndarray:
    """raw (S, 3, C, 2) -> X (S*C, 8)."""
    raw = np.asarray(raw, dtype=np.float64)
    S, C = raw.shape[0], tri_size(N)
    deg, ordm = tri_tables(N)
    sc = np.asarray(in_scale, dtype=np.float64).reshape(3, N + 1)[:, deg]  # (3, C)
    v = raw / sc[None, :, :, None]
    X = np.empty((S, C, FIN))
    X[:, :, 0:6] = v.transpose(0, 2, 1, 3).reshape(S, C, 6)
    X[:, :, 6] = deg / N
    X[:, :, 7] = ordm / N
    return X.reshape(S * C, FIN)


def targets(raw, out_scale, N: int) -> np.ndarray:
    raw = np.asarray(raw, dtype=np.float64)
    S, C = raw.shape[0], tri_size(N)
    deg, _ = tri_tables(N)
    sc = np.asarray(out_scale, dtype=np.float64).reshape(3, N + 1)[:, deg]
    v = raw / sc[None, :, :, None]
    return v.transpose(0, 2, 1, 3).reshape(S * C, 6)


def _pool(H, S, N):
    """mean over m <= n of rows of degree n: (S*C, F) -> (S, N+1, F)"""
    deg, _ = tri_tables(N)
    C = tri_size(N)
    Hs = H.reshape(S, C, -1)
    P = np.zeros((S, N + 1, H.shape[1]))
    np.add.at(P, (slice(None), deg), Hs)
    return P / (np.arange(N + 1) + 1.0)[None, :, None]


def _degsum(D, S, N):
    deg, _ = tri_tables(N)
    C = tri_size(N)
    Q = np.zeros((S, N + 1, D.shape[1]))
    np.add.at(Q, (slice(None), deg), D.reshape(S, C, -1))
    return Q


def _unpool(G, S, N):
    """(S, N+1, F) -> rows: G[s, deg] / (deg + 1)"""
    deg, _ = tri_tables(N)
    return (G[:, deg, :] / (deg + 1.0)[None, :, None]).reshape(S * tri_size(N), -1)


def forward(p: dict, X, S: int, N: int) -> dict:
    deg, _ = tri_tables(N)
    f = {"X": X}
    f["H1"] = np.tanh(X @ p["W1"].T + p["b1"])
    f["P1"] = _pool(f["H1"], S, N)
    mix = (f["P1"] @ p["V2"].T)[:, deg, :].reshape(S * tri_size(N), -1)
    f["H2"] = np.tanh(f["H1"] @ p["W2"].T + p["b2"] + mix)
    f["P2"] = _pool(f["H2"], S, N)
    mix = (f["P2"] @ p["V3"].T)[:, deg, :].reshape(S * tri_size(N), -1)
    f["H3"] = np.tanh(f["H2"] @ p["W3"].T + p["b3"] + mix)
    f["Y"] = f["H3"] @ p["W4"].T + p["b4"]
    return f


def loss_grad(params, W: int, N: int, in_scale, out_scale, x, target, round_weights: bool = True):
    """Loss and gradient (flat, in the parameter layout) on samples x (S, 3, C, 2) against target corrections."""
    p = unpack(params, W, round_weights)
    S = np.asarray(x).shape[0]
    X = features(x, in_scale, N)
    T = targets(target, out_scale, N)
    f = forward(p, X, S, N)
    mask = np.tile(trained_mask(N), (S, 1))
    count = float(mask.sum())
    d = np.where(mask, f["Y"] - T, 0.0)
    loss = float((d * d).sum() / count)
    dY = 2.0 * d / count
    g = {"W4": dY.T @ f["H3"], "b4": dY.sum(0)}
    dH3 = dY @ p["W4"]

    def inner(dH, Hout, Hin, Pin, Wn, bn, Vn):
        dZ = dH * (1.0 - Hout * Hout)
        g[Wn] = dZ.T @ Hin
        g[bn] = dZ.sum(0)
        Q = _degsum(dZ, S, N)
        g[Vn] = Q.reshape(-1, W).T @ Pin.reshape(-1, W)
        return dZ @ p[Wn] + _unpool(Q @ p[Vn], S, N)

    dH2 = inner(dH3, f["H3"], f["H2"], f["P2"], "W3", "b3", "V3")
    dH1 = inner(dH2, f["H2"], f["H1"], f["P1"], "W2", "b2", "V2")
    dZ1 = dH1 * (1.0 - f["H1"] ** 2)
    g["W1"] = dZ1.T @ X
    g["b1"] = dZ1.sum(0)
    return loss, pack(g, W)


def correction(params, W: int, N: int, in_scale, out_scale, x):
    """The correction (S, 3, C, 2) in raw units and the per-sample rms of the trained outputs."""
    p = unpack(params, W, True)
    S = np.asarray(x).shape[0]
    C = tri_size(N)
    f = forward(p, features(x, in_scale, N), S, N)
    mask = trained_mask(N)
    Y = np.where(mask[None], f["Y"].reshape(S, C, 6), 0.0)
    deg, _ = tri_tables(N)
    sc = np.asarray(out_scale, dtype=np.float64).reshape(3, N + 1)[:, deg]  # (3, C)
    corr = Y.reshape(S, C, 3, 2).transpose(0, 2, 1, 3) * sc[None, :, :, None]
    res = np.sqrt((Y * Y).sum(axis=(1, 2)) / mask.sum())
    return corr, res


def adam(params, grad, m1, m2, t: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """One Adam update (t = 1 for the first), returning new (params, m1, m2)."""
    m1 = b1 * m1 + (1 - b1) * grad
    m2 = b2 * m2 + (1 - b2) * grad * grad
    mh = m1 / (1 - b1 ** t)
    vh = m2 / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m1, m2
