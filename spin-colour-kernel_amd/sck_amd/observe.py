"""Observable series, accumulators and error analysis on top of the FP64 reduction kernels.

* `series`: energy and |M| per replica after every `every` sweeps, captured with the sweeps in one HIP graph
  (the reference captures "sweep + observe" the same way, SCK/simulate.py:361-376).
* `accumulate`: the reference's per-temperature accumulators about a reference energy E_ref (dE, dE^2, |M|,
  |M|^2; SCK/simulate.py:349-398) and Cv = var(E) / (T^2 N), chi = N var(|M|) / T per replica, averaged
  (SCK/simulate.py:521-526).
* `integrated_tau`: Sokal's automatic window, c = 6, ACF averaged over chains (SCK/exp_core.py:157-173).
* `jackknife`: delete-one-block jackknife error of any statistic of per-chain data (new: the reference
  jackknifes only Tc, remote/e4_jackknife.py:63-80).
* `effective_samples`: sum over temperatures of chains x measurements / (2 tau_int).
"""

from __future__ import annotations

import time

import numpy as np
import torch

__all__ = ["series", "accumulate", "integrated_tau", "jackknife", "effective_samples", "cv_chi"]


def series(sampler, S: torch.Tensor, n_meas: int, every: int = 1, fused: bool = False, anchor_every: int = 100,
           log=None):
    """(E (n_meas, B), |M|/N (n_meas, B), seconds) after each of n_meas replays of a graph of `every` sweeps;
    E in float64 totals, |M| the norm of the total magnetization per active site.

    fused=False: the graph ends with one observation (the float64 reductions over the lattice). fused=True (tiled
    samplers): the sweeps keep running totals up to date (Sampler.track), re-anchored by an observation every
    `anchor_every` measurements, so a measurement costs no pass over the lattice. With the tiled sweep and an odd
    `every`, two graphs with the spins and the workspace exchanged are replayed in turn (Sampler.graph_pair), so
    no measurement copies the lattice back; S holds the final state on return. `log`, if given, receives a
    progress line about every tenth of the measurements."""
    B = sampler.B
    N = float(sampler.sys.n_active)
    dev = sampler.device
    if fused:
        sampler.track(S)
    there, back, W = sampler.graph_pair(S, every, observe=not fused)
    obs = sampler._track if fused else sampler._obs
    nv = sampler.sys.b * sampler.sys.n
    out = torch.zeros((n_meas, B * (1 + nv)), dtype=torch.float64, device=dev)
    there()                                    # warm-up: both replays, the state back in S
    back()
    if fused:
        sampler.track(S)
    torch.cuda.synchronize(dev)
    step = max(1, n_meas // 10)
    t0 = time.perf_counter()
    for i in range(n_meas):
        cur = W if i % 2 else S                # the buffer holding the state before this replay
        if fused and i and i % max(1, int(anchor_every)) == 0:
            sampler.track(cur)
        (back if i % 2 else there)()
        out[i].copy_(obs)
        if log is not None and (i + 1) % step == 0:
            log(f"    {i + 1}/{n_meas} measurements, {time.perf_counter() - t0:.1f} s")
    if n_meas % 2 and W is not S:
        S.copy_(W)
    torch.cuda.synchronize(dev)
    if fused:
        sampler.untrack()
    dt = time.perf_counter() - t0
    E = out[:, :B]
    M = out[:, B:].reshape(n_meas, B, sampler.sys.b, sampler.sys.n).sum(dim=2)
    return E.cpu().numpy(), (M.norm(dim=2) / N).cpu().numpy(), dt


def integrated_tau(x: np.ndarray, c: float = 6.0) -> float:
    """Integrated autocorrelation time of x (time, chains): tau = 1 + 2 sum_{t <= M} rho(t), the window M the
    first with M >= c tau; the ACF is averaged over chains via FFT; a constant series has tau = 1."""
    x = np.asarray(x, np.float64)
    if x.ndim == 1:
        x = x[:, None]
    x = x - x.mean(axis=0)
    n = x.shape[0]
    if not np.all(np.isfinite(x)) or float((x * x).sum()) == 0.0:
        return 1.0
    f = np.fft.rfft(x, n=2 * n, axis=0)
    acf = np.fft.irfft(f * np.conj(f), axis=0)[:n].mean(axis=1)
    acf = acf / acf[0]
    if not np.isfinite(acf).all():
        return 1.0
    tau = 1.0
    csum = np.cumsum(acf[1:])
    for M in range(1, n):
        tau = 1.0 + 2.0 * csum[M - 1]
        if M >= c * tau:
            break
    return max(float(tau), 0.5)


def jackknife(stat, data: np.ndarray, blocks: int = 16):
    """(value, error) of stat(data) with delete-one-block jackknife over the last axis (chains)."""
    data = np.asarray(data)
    n = data.shape[-1]
    blocks = max(2, min(blocks, n))
    edges = np.linspace(0, n, blocks + 1).astype(int)
    full = stat(data)
    vals = []
    for b in range(blocks):
        keep = np.concatenate([np.arange(0, edges[b]), np.arange(edges[b + 1], n)])
        vals.append(stat(data[..., keep]))
    vals = np.asarray(vals)
    err = np.sqrt((blocks - 1) / blocks * ((vals - vals.mean(axis=0)) ** 2).sum(axis=0))
    return full, err


def cv_chi(E: np.ndarray, Mabs: np.ndarray, T: float, N: float, pooled: bool = False):
    """Cv = var(E)/(T^2 N) and chi = N var(|M|)/T from series (time, chains) with jackknife errors over chains.

    pooled=False: the variance of each chain's own series, averaged over chains (the reference's estimator,
    SCK/simulate.py:521-526); it is biased low by about 2 tau_int / n for a series of n correlated samples.
    pooled=True: one variance over every (time, chain) sample of chains that sample the same equilibrium,
    free of that tau-dependent bias (used to compare proposals with different tau)."""
    def var(x):
        return x.reshape(-1).var() if pooled else x.var(axis=0).mean()

    def cv(e):
        return float(var(e) / (T * T * N))

    def chi(m):
        return float(N * var(m) / T)
    return jackknife(cv, E), jackknife(chi, Mabs)


def accumulate(sampler, S: torch.Tensor, temps, n_therm: int, n_meas: int, every: int = 1):
    """Temperature ladder with the reference's accumulators; returns per temperature E, E2, |M|, M2 per replica
    and Cv, chi per site averaged over replicas."""
    temps = np.asarray(temps, float)
    N = float(sampler.sys.n_active)
    out = {k: np.zeros((len(temps), sampler.B)) for k in ("E", "E2", "Mabs", "M2")}
    cv, chi = np.zeros(len(temps)), np.zeros(len(temps))
    for t, T in enumerate(temps):
        sampler.set_temperature(float(T))
        sampler.sweep(S, int(n_therm))
        E, Ma, _ = series(sampler, S, int(n_meas), every)
        Er = E[0]
        dE = E - Er
        out["E"][t] = Er + dE.mean(0)
        out["E2"][t] = (dE * dE).mean(0) + 2 * Er * dE.mean(0) + Er * Er
        out["Mabs"][t] = Ma.mean(0)
        out["M2"][t] = (Ma * Ma).mean(0)
        varE = np.maximum(out["E2"][t] - out["E"][t] ** 2, 0.0)
        varM = np.maximum(out["M2"][t] - out["Mabs"][t] ** 2, 0.0)
        cv[t] = (varE / (T * T * N)).mean()
        chi[t] = (N * varM / T).mean()
    out["Cv"], out["chi"] = cv, chi
    return out


def effective_samples(x: np.ndarray, groups: np.ndarray):
    """Effective independent samples of series x (time, B) whose replicas fall into temperature groups:
    sum over groups of chains * time / (2 tau_int of the group), and the per-group tau."""
    groups = np.asarray(groups)
    taus = {}
    ess = 0.0
    for g in np.unique(groups):
        xs = x[:, groups == g]
        tau = integrated_tau(xs)
        taus[int(g)] = tau
        ess += xs.shape[1] * xs.shape[0] / (2.0 * tau)
    return ess, taus
