"""Effective independent samples per second: learned cone width against fixed-width Metropolis, same GPU.

Three proposals run from one common equilibrium state (equilibrated with the learned width, the fastest-mixing
chain; every proposal keeps detailed balance, so the state stays stationary under each): the uniform redraw
(the reference's proposal, SCK/generator.py:281-289), a scan of fixed half-angles, and the learned width. The
fixed-width baseline is the scanned width with the best measured ESS/s (geometric mean over the ladder), a
tuned fixed-width Metropolis. tau_int comes from Sokal's window on the energy and |M| series
of each temperature (sck_amd.observe). Reported gains of one proposal over another: per temperature, the
ratio of ESS/s; their geometric mean over the ladder (the headline); the bottleneck ratio (slowest
temperature against slowest temperature); and the ratio of the summed ESS/s, which the fastest (hottest)
temperatures dominate. Observables E, |M|, Cv and chi carry jackknife errors so the proposals are checked
against each other within 2 sigma.
"""

from __future__ import annotations

import time

import numpy as np
import torch

from .observe import cv_chi, effective_samples, jackknife, series
from .proposal import LearnedWidth, best_fixed_width, train_width
from .sampler import Sampler

__all__ = ["compare_proposals", "summarize_proposals"]


def _stats(E, Ma, temps_rep, temps, N):
    out = []
    for j, T in enumerate(temps):
        sel = temps_rep == j
        e, m = E[:, sel], Ma[:, sel]
        Em, Ee = jackknife(lambda x: x.mean() / N, e)
        Mm, Me = jackknife(lambda x: x.mean(), m)
        # pooled variances: the per-chain estimator's bias grows with tau, which differs between proposals
        (cv, cve), (chi, chie) = cv_chi(e, m, float(T), N, pooled=True)
        out.append(dict(T=float(T), E=float(Em), E_err=float(Ee), M=float(Mm), M_err=float(Me), Cv=cv, Cv_err=float(cve),
                        chi=chi, chi_err=float(chie)))
    return out


def _gains(a: dict, b: dict, ntemp: int) -> dict:
    """ESS/s of proposal b over proposal a: per temperature (tau_a / tau_b) x (time_a / time_b), their
    geometric mean over the ladder, the bottleneck ratio and the ratio of the summed ESS/s."""
    out = {}
    for key in ("E", "M"):
        r = np.array([a[f"tau_{key}"][j] / b[f"tau_{key}"][j] * a["seconds"] / b["seconds"] for j in range(ntemp)])
        out[f"gain_{key}_per_T"] = r.tolist()
        out[f"gain_{key}"] = float(np.exp(np.log(r).mean()))
        ta = np.array([a[f"tau_{key}"][j] for j in range(ntemp)])
        tb = np.array([b[f"tau_{key}"][j] for j in range(ntemp)])
        out[f"gain_{key}_bottleneck"] = float(ta.max() / tb.max() * a["seconds"] / b["seconds"])
        out[f"gain_{key}_sum"] = b[f"ess_{key}_per_s"] / a[f"ess_{key}_per_s"]
    return out


def _zscores(a: dict, b: dict) -> np.ndarray:
    z = []
    for x, y in zip(a["observables"], b["observables"]):
        for k in ("E", "M", "Cv", "chi"):
            err = np.hypot(x[k + "_err"], y[k + "_err"])
            z.append(abs(x[k] - y[k]) / err if err > 0 else 0.0)
    return np.asarray(z)


def compare_proposals(system, B: int, temps, warmup: int = 200, equil: int = 1000, therm: int = 50,
                      n_meas: int = 1000, every: int = 1, seed: int = 0, train_steps: int = 400, log=print,
                      widths=(0.5, 1.0, 1.5, 2.0, 2.5), init=None, fine_table: bool = True) -> dict:
    """Train the width on `warmup` fixed-width sweeps, equilibrate one state for `equil` sweeps with the
    learned width, then from that common equilibrium state run `therm` + `n_meas` sweeps of each of: the
    uniform redraw (the reference's proposal), each fixed width of `widths`, and the learned width. The
    fixed baseline is the width with the largest geometric-mean ESS/s of E over the ladder, measured. With
    fine_table, the learned network also runs from a table 4x finer along every feature: the geometric mean over
    the ladder of tau_E(table) / tau_E(fine table) is the table's cost in effective samples per sweep."""
    temps = np.asarray(temps, np.float32)
    per = B // len(temps)
    T_rep = np.repeat(temps, per)
    groups = np.repeat(np.arange(len(temps)), per)
    N = float(system.n_active)
    # warm-up with a moderate fixed width; its records train the network and pick the fixed baseline
    W = Sampler(system, B, "cuda", seed=seed, proposal="cone", width=1.0)
    W.set_temperature(torch.as_tensor(T_rep))
    # init(sampler) -> spins: e.g. drawn on the device for lattices too large for the host generator
    S0 = W.init_spins(seed=seed) if init is None else init(W)
    log(f"  initial state ready ({B} replicas x {int(N)} sites); warm-up and training")
    t0 = time.perf_counter()
    lw = train_width(W, S0, warmup=warmup, steps=train_steps, log=log)
    t_train = time.perf_counter() - t0
    w_fixed, scan = best_fixed_width(W, lw.samples)
    del W                                      # its spins and workspace are no longer needed
    torch.cuda.empty_cache()
    log(f"  fixed-width scan (width, mean log expected squared jump): "
        f"{[(w, round(j, 5)) for w, j in scan]} -> {w_fixed}")
    result = dict(system=system.name, B=B, temps=temps.tolist(), warmup=warmup, equil=equil, therm=therm,
                  n_meas=n_meas, every=every, fixed_width=w_fixed, width_scan=scan, train_seconds=t_train,
                  train_history=lw.history)
    # one equilibrium state for all proposals: any chain that keeps detailed balance leaves it stationary
    Q = Sampler(system, B, "cuda", seed=seed + 7, proposal="learned", learned=lw)
    Q.set_temperature(torch.as_tensor(T_rep))
    state, S0 = S0, None                       # the warm-up state itself: nothing else reads it
    Q.sweep(state, equil)
    del Q
    torch.cuda.empty_cache()
    # the same network tabulated 4x finer along every feature: what the sweep's table costs against the
    # network it stands for
    lw_fine = LearnedWidth(lw.params, dict(lw.cfg), n=tuple(4 * int(x) for x in lw.n),
                           s=tuple(4.0 * float(x) for x in lw.s))
    kinds = [("uniform", dict(proposal="uniform"))] + \
        [(f"fixed_{w:g}", dict(proposal="cone", width=float(w))) for w in widths] + \
        [("learned", dict(proposal="learned", learned=lw))] + \
        ([("learned_fine", dict(proposal="learned", learned=lw_fine))] if fine_table else [])
    for label, kw in kinds:
        A = Sampler(system, B, "cuda", seed=seed + 1, **kw)
        A.set_temperature(torch.as_tensor(T_rep))
        S = state.clone()
        A.sweep(S, therm)
        # the tiled sweep keeps running totals of E and M, so its measurements need no pass over the lattice
        fused = bool(getattr(A, "use_tiled", False)) and A.module.track_fold is not None
        E, Ma, secs = series(A, S, n_meas, every, fused=fused, log=log)
        essE, tauE = effective_samples(E, groups)
        essM, tauM = effective_samples(Ma, groups)
        del A, S
        torch.cuda.empty_cache()
        result[label] = dict(seconds=secs, sweeps=n_meas * every, fused=fused, ess_E=essE, ess_M=essM,
                             ess_E_per_s=essE / secs, ess_M_per_s=essM / secs, tau_E=tauE, tau_M=tauM,
                             updates_per_s=B * N * n_meas * every / secs,
                             observables=_stats(E, Ma, groups, temps, N))
        log(f"  {label:8s}: {secs:7.2f} s, ESS/s (E) {essE / secs:10.1f}, ESS/s (|M|) {essM / secs:10.1f}, "
            f"updates/s {B * N * n_meas * every / secs:.3e}")
    summarize_proposals(result, len(temps), widths, fine_table)
    log(f"  best fixed width by measured ESS/s: {result['fixed_width']}")
    return result


def summarize_proposals(result: dict, nt: int, widths, fine_table: bool = True) -> dict:
    """The comparison's figures from the per-proposal runs in `result` (keys "uniform", "fixed_<w>" for w in
    widths, "learned" and, with fine_table, "learned_fine"; each with seconds, tau_E and tau_M per temperature
    group, ess_E_per_s, ess_M_per_s and observables): the best fixed width by geometric-mean ESS/s of E, the
    learned width's gains over it and over the uniform redraw, the per-temperature ceiling of constant widths,
    the table's cost against the finer table, and the agreement of the observables. Updates `result`."""
    # the fixed baseline: the measured-best constant width (geometric mean over the ladder of ESS/s of E)
    def gm(r):
        return float(np.exp(np.mean([np.log(1.0 / (r["tau_E"][j] * r["seconds"])) for j in range(nt)])))
    fixed_labels = [f"fixed_{w:g}" for w in widths]
    best = max(fixed_labels, key=lambda k: gm(result[k]))
    result["fixed"] = result[best]
    # the width the training objective picked, replaced by the one measured best
    result["fixed_width_by_objective"] = result.get("fixed_width")
    result["fixed_width"] = float(best.split("_", 1)[1])
    result["fixed_scan_gm_ess"] = {k: gm(result[k]) for k in fixed_labels + ["uniform", "learned"]}
    result.update(_gains(result["fixed"], result["learned"], nt))
    # the ceiling of any width-only proposal that is constant per temperature: at every temperature the best
    # of the scanned widths and the uniform redraw, each by its own measured ESS/s of E
    cands = fixed_labels + ["uniform"]
    best_T = [max(cands, key=lambda k: 1.0 / (result[k]["tau_E"][j] * result[k]["seconds"])) for j in range(nt)]
    ceil = np.array([1.0 / (result[k]["tau_E"][j] * result[k]["seconds"]) for j, k in enumerate(best_T)])
    lrn = np.array([1.0 / (result["learned"]["tau_E"][j] * result["learned"]["seconds"]) for j in range(nt)])
    result["per_T_best_fixed"] = best_T
    result["gain_E_per_T_best_over_fixed"] = float(np.exp(np.mean(np.log(
        ceil * np.array([result["fixed"]["tau_E"][j] * result["fixed"]["seconds"] for j in range(nt)])))))
    result["gain_E_vs_per_T_best"] = float(np.exp(np.mean(np.log(lrn / ceil))))
    if fine_table:
        result["table_vs_fine_tau_E"] = float(np.exp(np.mean([
            np.log(result["learned"]["tau_E"][j] / result["learned_fine"]["tau_E"][j]) for j in range(nt)])))
    result["vs_uniform"] = _gains(result["uniform"], result["learned"], nt)
    result["fixed_vs_uniform"] = _gains(result["uniform"], result["fixed"], nt)
    # agreement of the observables: learned against each baseline
    z = np.concatenate([_zscores(result["fixed"], result["learned"]), _zscores(result["uniform"], result["learned"])])
    result["max_z"] = float(np.max(z))
    result["frac_within_2sigma"] = float(np.mean(z < 2.0))
    return result
