"""CPU oracle of the learned-proposal objective and its gradient, float64 (test infrastructure only).

Restates spin-colour-kernel_amd/csrc/mlp.hip in numpy: the features of a training sample, the 3 -> 32 -> 32 -> 1 tanh
network, the cone proposal of half-angle w = wmin + (wmax - wmin) sigmoid(o) with draws (u1, u2) from Philox
(key = seed, counter = sample * ndraw + draw), a sample's expected squared jump J_i = mean over its draws of
min(1, exp(-beta dE)) |t - s|^2, and the gradient of L = -mean_i log(1e-4 + J_i) by backpropagation. The reference has no learned proposal, so this
part is pinned by mathematics (finite differences of the same objective) rather than by reference vectors.
"""

from __future__ import annotations

import numpy as np

from . import oracle as O

H = 32
NIN = 3
NPARAM = H * NIN + H + H * H + H + H + 1
OFF_W1, OFF_B1 = 0, H * NIN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H


def unpack_params(P):
    P = np.asarray(P, np.float64)
    return (P[OFF_W1:OFF_B1].reshape(H, NIN), P[OFF_B1:OFF_W2], P[OFF_W2:OFF_B2].reshape(H, H), P[OFF_B2:OFF_W3],
            P[OFF_W3:OFF_B3], P[OFF_B3])


def features(samples: np.ndarray, cfg: dict) -> np.ndarray:
    """samples: (N, 12) rows (s[3], v[3], beta, k1, ku, axis[3]) -> (N, 3) scaled features."""
    S = np.asarray(samples, np.float64)
    n = cfg["n"]
    v = S[:, 3:3 + n]
    hv = np.sqrt((v * v).sum(1))
    e = np.asarray(cfg["axis"], np.float64)[:n]
    f0 = hv * S[:, 6]
    f1 = (abs(cfg["du"]) + np.abs(S[:, 8]) + np.abs(S[:, 7])) * S[:, 6]
    f2 = np.where(hv > 0, np.abs(v @ e) / np.where(hv > 0, hv, 1.0), 0.0)
    return np.stack([f0, f1, f2], 1) * np.asarray(cfg["xs"], np.float64)[None, :]


def forward(P, x):
    W1, b1, W2, b2, W3, b3 = unpack_params(P)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3 + b3, h1, h2


def draws(nsample: int, ndraw: int, seed: int):
    ids = np.arange(nsample * ndraw, dtype=np.uint32)
    ctr = np.stack([ids, np.zeros_like(ids), np.zeros_like(ids), np.zeros_like(ids)], 1)
    raw = O.philox((int(seed) & 0xFFFFFFFF, 0), ctr)
    u = np.array([[O.u01(int(a)), O.u01(int(b))] for a, b, _, _ in raw], np.float64)
    return u[:, 0].reshape(nsample, ndraw), u[:, 1].reshape(nsample, ndraw)


def _frame(s):
    nz = np.abs(s[:, 2]) > 0.9
    ax = np.where(nz[:, None], np.array([1.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0]))
    e1 = np.cross(ax, s)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(s, e1)
    return e1, e2


def cone(s, w, u1, u2, n):
    """Trial spins and dt/dw (N, n) for half-angles w (N,)."""
    if n == 2:
        r = 2 * u1 - 1
        ph = np.arctan2(s[:, 1], s[:, 0]) + w * r
        t = np.stack([np.cos(ph), np.sin(ph)], 1)
        dt = np.stack([-np.sin(ph) * r, np.cos(ph) * r], 1)
        return t, dt
    c = 1 - u1 * (1 - np.cos(w))
    dc = -u1 * np.sin(w)
    st = np.sqrt(np.maximum(1 - c * c, 1e-30))
    dst = -c * dc / st
    phi = 2 * np.pi * u2
    e1, e2 = _frame(s)
    perp = np.cos(phi)[:, None] * e1 + np.sin(phi)[:, None] * e2
    return c[:, None] * s + st[:, None] * perp, dc[:, None] * s + dst[:, None] * perp


def delta_e(S, t, cfg):
    """dE(s -> t) and d dE / dt for sample rows S."""
    n = cfg["n"]
    s, v = S[:, :n], S[:, 3:3 + n]
    dE = -((t - s) * v).sum(1)
    g = -v.copy()
    k1 = S[:, 7]
    if n == 3:
        k2 = cfg["k2"]

        def cub(p):
            x, y, z = p[:, 0], p[:, 1], p[:, 2]
            return k1 * (x * x * y * y + x * x * z * z + y * y * z * z) + k2 * x * x * y * y * z * z
        dE = dE + cub(t) - cub(s)
        x, y, z = t[:, 0], t[:, 1], t[:, 2]
        g = g + np.stack([2 * k1 * x * (y * y + z * z) + 2 * k2 * x * y * y * z * z,
                          2 * k1 * y * (x * x + z * z) + 2 * k2 * y * x * x * z * z,
                          2 * k1 * z * (x * x + y * y) + 2 * k2 * z * x * x * y * y], 1)
    e = np.asarray(cfg["axis"], np.float64)[:n]
    du = cfg["du"]
    if du != 0.0:
        te, se = t @ e, s @ e
        dE = dE + du * (se * se - te * te)
        g = g - 2 * du * te[:, None] * e[None, :]
    ku, ax = S[:, 8], S[:, 9:9 + n]
    tn, sn = (t * ax).sum(1), (s * ax).sum(1)
    dE = dE + ku * (sn * sn - tn * tn)
    g = g - 2 * (ku * tn)[:, None] * ax
    return dE, g


EPS_J = 1e-4


def objective_and_grad(P, samples, cfg):
    """(mean_i log(EPS_J + J_i), dL/dP) in float64 with J_i a sample's expected squared jump over its ndraw
    draws and L = -mean_i log(EPS_J + J_i)."""
    S = np.asarray(samples, np.float64)
    N, K, n = S.shape[0], cfg["ndraw"], cfg["n"]
    x = features(S, cfg)
    o, h1, h2 = forward(P, x)
    sg = 1 / (1 + np.exp(-o))
    w = cfg["wmin"] + (cfg["wmax"] - cfg["wmin"]) * sg
    dwdo = (cfg["wmax"] - cfg["wmin"]) * sg * (1 - sg)
    u1, u2 = draws(N, K, cfg["seed"])
    Ji = np.zeros(N)
    dJdw = np.zeros(N)
    for k in range(K):
        t, dt = cone(S[:, :n], w, u1[:, k], u2[:, k], n)
        dE, g = delta_e(S, t, cfg)
        xb = -S[:, 6] * dE
        a = np.where(xb >= 0, 1.0, np.exp(np.minimum(xb, 0.0)))
        dadE = np.where(xb >= 0, 0.0, -S[:, 6] * a)
        d = ((t - S[:, :n]) ** 2).sum(1)
        Ji += a * d / K
        dJdw += ((d * dadE)[:, None] * g * dt + (a[:, None] * 2 * (t - S[:, :n])) * dt).sum(1) / K
    dLdo = -dJdw / (EPS_J + Ji) * dwdo / N
    obj = float(np.log(EPS_J + Ji).mean())
    W1, b1, W2, b2, W3, b3 = unpack_params(P)
    d2 = dLdo[:, None] * W3[None, :] * (1 - h2 * h2)
    d1 = (d2 @ W2) * (1 - h1 * h1)
    grad = np.concatenate([(d1.T @ x).reshape(-1), d1.sum(0), (d2.T @ h1).reshape(-1), d2.sum(0),
                           dLdo @ h2, [dLdo.sum()]])
    return obj, grad


def loss(P, samples, cfg) -> float:
    """L = -mean_i log(EPS_J + J_i), the quantity objective_and_grad differentiates."""
    S = np.asarray(samples, np.float64)
    N, K, n = S.shape[0], cfg["ndraw"], cfg["n"]
    o, _, _ = forward(P, features(S, cfg))
    w = cfg["wmin"] + (cfg["wmax"] - cfg["wmin"]) / (1 + np.exp(-o))
    u1, u2 = draws(N, K, cfg["seed"])
    Ji = np.zeros(N)
    for k in range(K):
        t, _ = cone(S[:, :n], w, u1[:, k], u2[:, k], n)
        dE, _ = delta_e(S, t, cfg)
        a = np.exp(np.minimum(-S[:, 6] * dE, 0.0))
        Ji += a * ((t - S[:, :n]) ** 2).sum(1) / K
    return float(-np.log(EPS_J + Ji).mean())
