"""Distribution-level validation: sampled energies against the exact Boltzmann distribution of a small Ising
system (the reference's experiment E2, SCK/exp_core.py:137-265).

`exact_energy_distribution` enumerates every state of a system of up to 24 active sites (bit-sliced over
numpy, any crystal, periodic or open axes, masks, field). `energy_test` samples one chain batch per seed,
thins by 2 tau_int from a pilot series, merges bins of expected count below five, and applies a chi-square
test per seed with Holm's step-down correction over the seeds. The coloured sweep must pass; the synchronous
all-sites control, which does not sample the Boltzmann distribution, must be rejected.
"""

from __future__ import annotations

import math

import numpy as np

from .observe import integrated_tau
from .system import System

__all__ = ["exact_energy_distribution", "merge_bins", "holm", "energy_test"]

MAX_SITES = 24


def _bond_pairs(s: System):
    """(i, j, J) for every (site, bond) with an active neighbour inside the box, i and j indices of active
    sites; each bond appears once from each end."""
    m = s.site_mask()
    idx = -np.ones(m.shape, np.int64)
    idx[m] = np.arange(int(m.sum()))
    L = (s.L1, s.L2, s.L3)
    out = []
    for a, z, y, x in zip(*np.nonzero(m)):
        c = (x, y, z)
        for bd in s.crystal.bonds:
            if bd.alpha != a:
                continue
            J = s.J.get(bd.shell, 0.0)
            if J == 0.0:
                continue
            d = list(bd.delta) + [0] * (3 - s.crystal.D)
            q = []
            ok = True
            for k in range(3):
                v = c[k] + d[k]
                if k < s.crystal.D and s.periodic[k]:
                    v %= L[k]
                elif not 0 <= v < L[k]:
                    ok = False
                q.append(v)
            if ok and m[bd.alpha2, q[2], q[1], q[0]]:
                out.append((int(idx[a, z, y, x]), int(idx[bd.alpha2, q[2], q[1], q[0]]), float(J)))
    return out


def exact_energy_distribution(s: System, T: float) -> dict:
    """{"levels": distinct energies (rounded to 1e-6), "degeneracy": states per level, "p": Boltzmann
    probabilities at temperature T} of an Ising system by enumeration of all 2^N states."""
    if s.n != 1:
        raise ValueError("exact enumeration is for Ising systems")
    N = s.n_active
    if N > MAX_SITES:
        raise ValueError(f"exact enumeration needs at most {MAX_SITES} active sites, got {N}")
    states = np.arange(2 ** N, dtype=np.int64)
    spin = lambda i: ((states >> i) & 1) * 2 - 1          # noqa: E731
    E = np.zeros(2 ** N)
    for i, j, J in _bond_pairs(s):
        E -= 0.5 * J * spin(i) * spin(j)
    if s.h is not None and s.h[0] != 0.0:
        for i in range(N):
            E -= float(s.h[0]) * spin(i)
    levels, counts = np.unique(np.round(E, 6), return_counts=True)
    w = counts * np.exp(-(levels - levels.min()) / float(T))
    return {"levels": levels, "degeneracy": counts, "p": w / w.sum()}


def merge_bins(obs: np.ndarray, expected: np.ndarray, min_expected: float = 5.0):
    """Merge consecutive bins until each holds an expected count of at least min_expected (the remainder joins
    the last bin)."""
    o2, e2 = [], []
    co, ce = 0.0, 0.0
    for o, e in zip(obs, expected):
        co += o
        ce += e
        if ce >= min_expected:
            o2.append(co)
            e2.append(ce)
            co, ce = 0.0, 0.0
    if ce > 0 and e2:
        o2[-1] += co
        e2[-1] += ce
    return np.asarray(o2), np.asarray(e2)


def holm(p) -> list[float]:
    """Holm's step-down adjusted p-values."""
    p = np.asarray(p, float)
    m = len(p)
    adj = np.zeros(m)
    run = 0.0
    for r, i in enumerate(np.argsort(p)):
        run = max(run, (m - r) * p[i])
        adj[i] = min(1.0, run)
    return adj.tolist()


def energy_test(make_sampler, s: System, T: float, seeds=(1, 2, 3), samples: int = 20000, batch: int = 1000,
                rule: str = "coloured", alpha: float = 0.01, therm: int = 500, pilot: int = 400) -> dict:
    """Chi-square test of sampled energies against the exact distribution, one chain batch per seed.

    make_sampler(seed) returns a Sampler of `s` with `batch` replicas; rule "coloured" steps with sweep(),
    "all sites" with sync_sweep()."""
    from scipy import stats
    ex = exact_energy_distribution(s, T)
    lv = ex["levels"]
    rows = []
    pooled = np.zeros(len(lv))
    for seed in seeds:
        A = make_sampler(seed)
        A.set_temperature(float(T))
        S = A.init_spins(seed=seed)
        step = A.sweep if rule == "coloured" else A.sync_sweep
        step(S, therm)
        pil = []
        for _ in range(pilot):
            step(S, 1)
            pil.append(A.energy(S).cpu().numpy())
        tau = integrated_tau(np.array(pil))
        thin = max(1, int(math.ceil(2 * tau)))
        Es = []
        for _ in range(int(math.ceil(samples / batch))):
            step(S, thin)
            Es.append(A.energy(S).cpu().numpy())
        Es = np.round(np.concatenate(Es)[:samples], 6)
        idx = np.clip(np.searchsorted(lv, Es), 0, len(lv) - 1)
        off = int((np.abs(lv[idx] - Es) > 1e-4).sum())
        obs = np.bincount(idx, minlength=len(lv)).astype(float)
        pooled += obs
        o2, e2 = merge_bins(obs, ex["p"] * obs.sum())
        chi2, p = stats.chisquare(o2, e2) if len(o2) > 1 else (0.0, 1.0)
        rows.append({"seed": int(seed), "tau": float(tau), "thin": thin, "chi2": float(chi2), "p": float(p),
                     "bins": int(len(o2)), "cohens_w": math.sqrt(float(chi2) / obs.sum()), "off_level": off})
    for r, pa in zip(rows, holm([r["p"] for r in rows])):
        r["p_holm"] = pa
    return {"system": s.name, "T": float(T), "rule": rule, "rows": rows,
            "rejected": int(sum(r["p_holm"] < alpha for r in rows)), "levels": lv.tolist(),
            "exact": ex["p"].tolist(), "sampled": (pooled / pooled.sum()).tolist()}
