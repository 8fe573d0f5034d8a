"""Cooling ladder of magnetic nanodots on the MI355X sweep (the counterpart of SCK/anneal.py:1-270).

One state is one parameter row (Jex2, Jex3, Jex4, Kan1, KanS, Hex, KDM) of the nanodot Hamiltonian
(`system.nanodot`, Jex = 1 sets the energy unit), one generated module and one batch of replicas. Every
replica follows the same ladder from T_start to T_end in kelvin (beta = 1 / (kB T), kB in meV/K): n_therm
sweeps at each temperature, then one measurement every `every` sweeps during n_meas sweeps, accumulated about
references taken at the start of the temperature so that the variances keep their precision:

    obs[t] = (<dE>, <dE^2>, <dM>, <dM^2>, <Mz>, <Q>, E_ref, M_ref)      shape (n_T, 8, B)

with E the total energy, M = |sum_i s_i| / N_active, Mz its z component and Q the Berg-Luescher charge of the
middle layer (SCK/anneal.py:9-12, 164-170). Thermalization runs as replays of a HIP graph of 10 `every` sweeps
and each measurement as one replay of a graph of `every` sweeps plus the float64 observation kernels; the
accumulators stay on the device.

    python -m sck_amd.anneal --config run.json --states 0,2 --out DIR
"""

from __future__ import annotations

import argparse
import json
import time
from pathlib import Path

import numpy as np
import torch

from . import system as sy
from .observables import topological_charge
from .sampler import Sampler

KB = 0.086173404     # meV / K
PARAM_NAMES = ("Jex2", "Jex3", "Jex4", "Kan1", "KanS", "Hex", "KDM")

__all__ = ["KB", "PARAM_NAMES", "ACC_ROWS", "build_state", "active_shells", "check_exact", "Observer",
           "anneal_state", "reduce_obs", "main"]


def build_state(params, Rd: float, L: int, angle_deg: float, shells=(1,)) -> sy.System:
    """The nanodot of one parameter row [Jex2, Jex3, Jex4, Kan1, KanS, Hex, KDM] (SCK/anneal.py:40-44):
    Jex = 1, no second cubic constant, field at angle_deg from the plane."""
    if len(params) != len(PARAM_NAMES):
        raise ValueError(f"a state has {len(PARAM_NAMES)} parameters {PARAM_NAMES}, got {len(params)}")
    J2, J3, J4, K1, KS, H, D = (float(v) for v in params)
    return sy.nanodot(Rd=Rd, L=L, Jex=1.0, Kan1=K1, Kan2=0.0, KanS=KS, Hex=H, KDM=D, gamma_deg=angle_deg,
                      shells=tuple(shells), J_shells={2: J2, 3: J3, 4: J4})


def active_shells(rows) -> list[int]:
    """Shells with a non-zero coupling in any state, plus shell 1: one colouring serves every state
    (SCK/anneal.py:47-54)."""
    used = {1}
    for r in rows:
        used.update(k for k, v in zip((2, 3, 4), list(r)[:3]) if float(v) != 0.0)
    return sorted(used)


def check_exact(s: sy.System, device, sweeps: int = 3, T: float = 0.5, seed: int = 1) -> int:
    """Spins that differ between the coloured sweep and the sequential chain after `sweeps` sweeps of two
    replicas (must be 0; SCK/anneal.py:57-66)."""
    A, R = Sampler(s, 2, device, seed), Sampler(s, 2, device, seed)
    A.set_temperature(T)
    R.set_temperature(T)
    S1 = A.init_spins(seed=seed)
    S2 = S1.clone()
    A.sweep(S1, sweeps)
    R.reference_sweep(S2, sweeps)
    return int((S1 != S2).sum())


class Observer:
    """(E float64 total energy, mean magnetization vector (B, 3), middle-layer Q) per replica. E and M come
    from the sampler's float64 reduction kernels, directly (`__call__`) or from the buffers a graph with
    observe=True has just written (`from_buffers`); Q from the spins."""

    def __init__(self, A: Sampler):
        s = A.sys
        if s.n != 3:
            raise ValueError("the anneal ladder measures Heisenberg spins (n = 3)")
        self.A, self.s = A, s
        self.n_act = float(s.n_active)
        self.mid = s.L3 // 2

    def _combine(self, S, E, M):
        Mv = M.sum(dim=1) / self.n_act
        return E, Mv, topological_charge(S, self.s, self.mid)

    def __call__(self, S: torch.Tensor):
        E, M = self.A.observe(S)
        return self._combine(S, E, M)

    def from_buffers(self, S: torch.Tensor):
        E, M = self.A.observed()
        return self._combine(S, E, M)


def anneal_state(s: sy.System, replicas: int, temps_K, n_therm: int, n_meas: int, every: int, seed: int,
                 device="cuda", log=print) -> dict:
    """The ladder of one state (SCK/anneal.py:143-218). Returns obs (n_T, 8, B), mvec (n_T, B, 3), snaps
    (n_T, L2, L1) = middle-layer s_z of replica 0 at the end of each temperature, seconds, sweeps_per_chain
    and flips_per_ns."""
    A = Sampler(s, int(replicas), device, seed)
    S = A.init_spins(seed=seed)
    B = int(replicas)
    observe = Observer(A)
    k = max(1, int(every))
    big = 10 * k
    m = max(1, int(n_meas) // k)
    temps_K = np.asarray(temps_K, dtype=float)
    nT = len(temps_K)
    obs = np.zeros((nT, 8, B))
    mvec = np.zeros((nT, B, 3), np.float32)
    snaps = np.zeros((nT, s.L2, s.L1), np.float32)
    f64 = dict(dtype=torch.float64, device=A.device)
    E_ref, M_ref = torch.zeros(B, **f64), torch.zeros(B, **f64)
    acc = torch.zeros(6, B, **f64)                            # dE, dE^2, dM, dM^2, Mz, Q
    accv = torch.zeros(B, 3, **f64)
    A.set_temperature(KB * float(temps_K[0]))
    rep_big = A.graph(S, big)
    rep_meas = A.graph(S, k, observe=True)

    def fold():
        E, Mv, Q = observe.from_buffers(S)
        dE = E - E_ref
        dM = Mv.norm(dim=1) - M_ref
        acc.add_(torch.stack([dE, dE * dE, dM, dM * dM, Mv[:, 2], Q]))
        accv.add_(Mv)

    t0 = time.perf_counter()
    for t, TK in enumerate(temps_K):
        A.set_temperature(KB * float(TK))                    # in place: the graphs read beta by address
        for _ in range(int(n_therm) // big):
            rep_big()
        if int(n_therm) % big:
            A.sweep(S, int(n_therm) % big)
        E, Mv, _ = observe(S)
        E_ref.copy_(E)
        M_ref.copy_(Mv.norm(dim=1))
        acc.zero_()
        accv.zero_()
        for _ in range(m):
            rep_meas()
            fold()
        obs[t, :6] = (acc / m).cpu().numpy()
        obs[t, 6] = E_ref.cpu().numpy()
        obs[t, 7] = M_ref.cpu().numpy()
        mvec[t] = (accv / m).float().cpu().numpy()
        snaps[t] = S[0, 2, 0, observe.mid].float().cpu().numpy()
        if t % max(1, nT // 10) == 0 or t == nT - 1:
            log(f"    T = {TK:7.3f} K   <E>/site = {(obs[t, 6].mean() + obs[t, 0].mean()) / observe.n_act:+.4f}"
                f"   |M| = {obs[t, 7].mean() + obs[t, 2].mean():.4f}   {time.perf_counter() - t0:7.1f} s")
    torch.cuda.synchronize(A.device)
    dt = time.perf_counter() - t0
    sweeps = nT * (int(n_therm) + m * k)
    return dict(obs=obs, mvec=mvec, snaps=snaps, seconds=dt, sweeps_per_chain=sweeps,
                flips_per_ns=B * observe.n_act * sweeps / dt / 1e9)


# the rows of the accumulator array (n_T, 8, B), in the file format's order
ACC_ROWS = ("dE", "dE2", "dM", "dM2", "Mz", "Q", "E_ref", "M_ref")


def reduce_obs(obs, temps, n_spins, kB: float = KB) -> dict:
    """Per-temperature scalars from the accumulators (SCK/anneal.py:221-234): E per spin and |M| (means over
    replicas, with their spreads), Mz, Q (with its spread), Cv = var(E) / (kB T^2 N) and chi = var(|M|) / (kB T).
    Each replica's variance is taken from its own offsets about its reference value, then averaged."""
    acc = np.asarray(obs, dtype=np.float64)
    if acc.ndim != 3 or acc.shape[1] != len(ACC_ROWS):
        raise ValueError(f"expected accumulators of shape (n_T, {len(ACC_ROWS)}, batch), got {acc.shape}")
    row = dict(zip(ACC_ROWS, acc.transpose(1, 0, 2)))
    T = np.asarray(temps, dtype=np.float64)
    kT = kB * T[:, None]

    def spread(first, second):
        """Variance from the first two moments of the offsets."""
        return row[second] - row[first] ** 2

    per_replica = {
        "E": (row["E_ref"] + row["dE"]) / n_spins,
        "M": row["M_ref"] + row["dM"],
        "Mz": row["Mz"],
        "Cv": spread("dE", "dE2") / (kT * T[:, None] * n_spins),
        "chi": spread("dM", "dM2") / kT,
        "Q": row["Q"],
    }
    out = {"T": np.asarray(temps)}
    for key, values in per_replica.items():
        out[key] = values.mean(axis=1)
        if key in ("E", "M", "Q"):
            out[key + "_sd"] = values.std(axis=1)
    return out


def _ladder(cfg: dict):
    """The temperatures of a run configuration and its split of each temperature's sweeps: two thirds
    thermalize, the rest measure."""
    temps = np.linspace(float(cfg["T_start"]), float(cfg["T_end"]), int(cfg["n_temps"]))
    total = int(cfg["sweeps"])
    return temps, total * 2 // 3, total - total * 2 // 3


def _selected(spec: str, count: int) -> list[int]:
    chosen = [int(v) for v in spec.split(",") if v.strip()]
    bad = [i for i in chosen if not 0 <= i < count]
    if bad:
        raise SystemExit(f"state indices must lie in [0, {count}); got {bad}")
    return chosen


def main(argv=None) -> None:
    """Process states of a run configuration on one GPU and write one state_<name>.npz each
    (SCK/anneal.py:237-266). Config keys: params (rows), names, T_start, T_end, n_temps, sweeps, Rd, layers,
    angle, replicas, every, seed."""
    ap = argparse.ArgumentParser(prog="python -m sck_amd.anneal")
    ap.add_argument("--config", type=Path, required=True)
    ap.add_argument("--states", required=True, help="comma list of state indices")
    ap.add_argument("--out", type=Path, required=True)
    ap.add_argument("--device", default="cuda:0")
    a = ap.parse_args(argv)
    cfg = json.loads(a.config.read_text())
    rows = cfg["params"]
    chosen = _selected(a.states, len(rows))
    a.out.mkdir(parents=True, exist_ok=True)
    temps, n_therm, n_meas = _ladder(cfg)
    shells = active_shells(rows)
    for i in chosen:
        name = cfg["names"][i]

        def say(msg: str, name=name) -> None:
            print(f"[{name}] {msg.strip()}", flush=True)

        s = build_state(rows[i], cfg["Rd"], cfg["layers"], cfg["angle"], shells)
        differ = check_exact(s, a.device)
        say(f"{s.colouring.n_colours} colours ({s.colouring.describe()}); coloured sweeps against the sequential "
            f"chain: {differ} spins differ")
        if differ:
            raise SystemExit(f"state {name}: the coloured sweep does not reproduce the sequential chain")
        r = anneal_state(s, int(cfg["replicas"]), temps, n_therm, n_meas, int(cfg["every"]), int(cfg["seed"]) + i,
                         a.device, log=say)
        np.savez_compressed(a.out / f"state_{name}.npz", params=np.asarray(rows[i], np.float32), obs=r["obs"],
                            M=np.transpose(r["mvec"], (1, 0, 2)), snaps=r["snaps"],
                            temperatures=temps.astype(np.float32), n_active=s.n_active, seconds=r["seconds"],
                            flips_per_ns=r["flips_per_ns"], colours=s.colouring.n_colours)
        say(f"finished in {r['seconds']:.1f} s at {r['flips_per_ns']:.2f} spin flips/ns")


if __name__ == "__main__":
    main()
