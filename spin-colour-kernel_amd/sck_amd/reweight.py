"""Multi-histogram reweighting of energy series from several temperatures (the counterpart of
SCK/exp_perf.py:754-795 `multi_histogram`).

Method: A. M. Ferrenberg and R. H. Swendsen, Phys. Rev. Lett. 63 (1989) 1195 (also known as WHAM). With
N_k (weighted) samples at inverse temperature b_k and all samples pooled, the free energies f_k = ln Z_k solve

    Z_k = sum_i w_i exp(-b_k x_i) / D_i,      D_i = sum_k N_k exp(-b_k x_i - f_k),

and the canonical average at any b is <O>_b = sum_i O_i p_i with p_i proportional to w_i exp(-b x_i) / D_i.
Everything is done in logarithms (logsumexp) in float64, with energies shifted by their weighted mean e0 so
the exponents stay small; f_0 is fixed at 0. The iteration starts from thermodynamic integration of <E> over
b (trapezoid rule) and stops when no f_k moves by more than `tol`.
"""

from __future__ import annotations

import numpy as np
import torch

__all__ = ["multi_histogram"]


def _f64(x, device) -> torch.Tensor:
    return torch.as_tensor(np.asarray(x, dtype=np.float64), dtype=torch.float64, device=device)


def multi_histogram(E, obs, temps, grid, weights=None, device="cpu", lnZ0=None, tol: float = 1e-10,
                    iters: int = 5000, return_lnZ: bool = False):
    """Canonical averages of each observable in `obs` at the temperatures `grid`.

    E: energies (temps, samples) measured at `temps`; obs: a list of arrays of E's shape; weights: per-sample
    multiplicities (0 drops a sample, as bootstrap resampling does), default 1; lnZ0: starting free energies.
    Returns a list with one array (len(grid),) per observable, and ln Z (relative to the first temperature,
    energies shifted by their weighted mean) when return_lnZ is set."""
    Et = _f64(E, device)
    if Et.ndim != 2:
        raise ValueError("E must have shape (temperatures, samples)")
    nT, nS = Et.shape
    b = 1.0 / _f64(temps, device)
    if b.numel() != nT:
        raise ValueError("one temperature per row of E")
    w = torch.ones_like(Et) if weights is None else _f64(weights, device)
    if w.shape != Et.shape:
        raise ValueError("weights must have E's shape")
    obs = [_f64(o, device) for o in obs]
    if any(o.shape != Et.shape for o in obs):
        raise ValueError("every observable must have E's shape")
    e0 = float((Et * w).sum() / w.sum())
    x = (Et - e0).reshape(-1)
    lw = torch.log(w.reshape(-1))                             # -inf drops a sample
    lnN = torch.log(w.sum(dim=1))
    if lnZ0 is None:
        Em = (Et * w).sum(dim=1) / w.sum(dim=1) - e0
        f = torch.zeros(nT, dtype=torch.float64, device=Et.device)
        # d ln Z / d b = -<E>: integrate along the ladder
        f[1:] = torch.cumsum(-(b[1:] - b[:-1]) * 0.5 * (Em[1:] + Em[:-1]), dim=0)
    else:
        f = _f64(lnZ0, device).clone()
    bx = b[:, None] * x[None, :]

    def log_denominator(f):
        return torch.logsumexp(lnN[:, None] - bx - f[:, None], dim=0)

    for _ in range(int(iters)):
        f_new = torch.logsumexp(lw[None, :] - bx - log_denominator(f)[None, :], dim=1)
        f_new = f_new - f_new[0]
        moved = float((f_new - f).abs().max())
        f = f_new
        if moved < tol:
            break
    lD = log_denominator(f)
    g = 1.0 / _f64(grid, device)
    p = torch.softmax(lw[None, :] - g[:, None] * x[None, :] - lD[None, :], dim=1)     # (grid, samples)
    out = [(p @ o.reshape(-1)).cpu().numpy() for o in obs]
    return (out, f.cpu().numpy()) if return_lnZ else out
