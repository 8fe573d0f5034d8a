"""Effective independent samples per second: learned cone width against the best fixed width, same GPU.

Both samplers start from the same warmed-up spins. The fixed width is the single half-angle with the largest
training objective on the warm-up records (the learned network's own objective restricted to one constant),
so the comparison is against a tuned fixed-width Metropolis. tau_int comes from Sokal's window on the energy
and |M| series of each temperature (sck_amd.observe). Reported gains: per temperature, ESS/s of the learned
width over the fixed one; their geometric mean over the ladder (the headline); the bottleneck ratio (slowest
temperature against slowest temperature); and the ratio of the summed ESS/s, which the fastest (hottest)
temperatures dominate. Observables E, |M|, Cv and chi carry jackknife errors so the two proposals are checked
against each other within 2 sigma.
"""

from __future__ import annotations

import time

import numpy as np
import torch

from .observe import cv_chi, effective_samples, jackknife, series
from .proposal import best_fixed_width, train_width
from .sampler import Sampler

__all__ = ["compare_proposals"]


def _stats(E, Ma, temps_rep, temps, N):
    out = []
    for j, T in enumerate(temps):
        sel = temps_rep == j
        e, m = E[:, sel], Ma[:, sel]
        Em, Ee = jackknife(lambda x: x.mean() / N, e)
        Mm, Me = jackknife(lambda x: x.mean(), m)
        (cv, cve), (chi, chie) = cv_chi(e, m, float(T), N)
        out.append(dict(T=float(T), E=float(Em), E_err=float(Ee), M=float(Mm), M_err=float(Me), Cv=cv, Cv_err=float(cve),
                        chi=chi, chi_err=float(chie)))
    return out


def compare_proposals(system, B: int, temps, warmup: int = 200, therm: int = 300, n_meas: int = 1000,
                      every: int = 1, seed: int = 0, train_steps: int = 400, log=print) -> dict:
    temps = np.asarray(temps, np.float32)
    per = B // len(temps)
    T_rep = np.repeat(temps, per)
    groups = np.repeat(np.arange(len(temps)), per)
    N = float(system.n_active)
    # warm-up with a moderate fixed width; its records train the network and pick the fixed baseline
    W = Sampler(system, B, "cuda", seed=seed, proposal="cone", width=1.0)
    W.set_temperature(torch.as_tensor(T_rep))
    S0 = W.init_spins(seed=seed)
    t0 = time.perf_counter()
    lw = train_width(W, S0, warmup=warmup, steps=train_steps, log=log)
    t_train = time.perf_counter() - t0
    w_fixed, scan = best_fixed_width(W, lw.samples)
    log(f"  fixed-width scan (width, mean squared jump): {[(w, round(j, 5)) for w, j in scan]} -> {w_fixed}")
    result = dict(system=system.name, B=B, temps=temps.tolist(), warmup=warmup, therm=therm, n_meas=n_meas,
                  every=every, fixed_width=w_fixed, width_scan=scan, train_seconds=t_train,
                  train_history=lw.history)
    state = S0.clone()
    for label, kw in (("fixed", dict(proposal="cone", width=w_fixed)), ("learned", dict(proposal="learned", learned=lw))):
        A = Sampler(system, B, "cuda", seed=seed + 1, **kw)
        A.set_temperature(torch.as_tensor(T_rep))
        S = state.clone()
        A.sweep(S, therm)
        E, Ma, secs = series(A, S, n_meas, every)
        essE, tauE = effective_samples(E, groups)
        essM, tauM = effective_samples(Ma, groups)
        result[label] = dict(seconds=secs, sweeps=n_meas * every, ess_E=essE, ess_M=essM,
                             ess_E_per_s=essE / secs, ess_M_per_s=essM / secs, tau_E=tauE, tau_M=tauM,
                             updates_per_s=B * N * n_meas * every / secs,
                             observables=_stats(E, Ma, groups, temps, N))
        log(f"  {label:8s}: {secs:7.2f} s, ESS/s (E) {essE / secs:10.1f}, ESS/s (|M|) {essM / secs:10.1f}, "
            f"updates/s {B * N * n_meas * every / secs:.3e}")
    f, l = result["fixed"], result["learned"]
    result["gain_E_sum"] = l["ess_E_per_s"] / f["ess_E_per_s"]
    result["gain_M_sum"] = l["ess_M_per_s"] / f["ess_M_per_s"]
    # per temperature: ESS/s ratio = (tau_fixed / tau_learned) x (time_fixed / time_learned)
    for key in ("E", "M"):
        r = np.array([f[f"tau_{key}"][j] / l[f"tau_{key}"][j] * f["seconds"] / l["seconds"] for j in range(len(temps))])
        result[f"gain_{key}_per_T"] = r.tolist()
        result[f"gain_{key}"] = float(np.exp(np.log(r).mean()))              # geometric mean over the ladder
        tf = np.array([f[f"tau_{key}"][j] for j in range(len(temps))])
        tl = np.array([l[f"tau_{key}"][j] for j in range(len(temps))])
        result[f"gain_{key}_bottleneck"] = float(tf.max() / tl.max() * f["seconds"] / l["seconds"])
    # agreement of the two proposals' observables (z-scores)
    z = []
    for a, b in zip(f["observables"], l["observables"]):
        for k in ("E", "M", "Cv", "chi"):
            err = np.hypot(a[k + "_err"], b[k + "_err"])
            z.append(abs(a[k] - b[k]) / err if err > 0 else 0.0)
    result["max_z"] = float(np.max(z))
    result["frac_within_2sigma"] = float(np.mean(np.asarray(z) < 2.0))
    return result
