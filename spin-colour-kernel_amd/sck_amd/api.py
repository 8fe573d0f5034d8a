"""The short interface: describe, check and run a simulation, then read its results (SCK/api.py:1-226).

    from sck_amd import api as sck
    sck.describe()                                # models, crystals, parameters, defaults
    sck.check(crystal="sc", box=[16, 16, 16])     # errors and warnings, without running
    res = sck.run(crystal="sc", box=[16, 16, 16], states=[dict(name="ferromagnet", J={1: 1.0})],
                  T_max=2.0, T_min=0.5, T_step=0.1)
    res.table()                                   # E, |M|, Cv and chi per state and temperature

`run` uses every visible MI355X: one GPU runs in this process; several run one process each
(`python -m sck_amd.simulate`), each with an equal block of the replicas of every state. There is no CPU path.
It writes the reference's output folder and returns its results.
"""

from __future__ import annotations

import json
import os
import subprocess
import sys
import threading
import time
from pathlib import Path

import numpy as np

from . import crystals as cr
from . import simulate as SIM
from .hamiltonian import MODELS, build_system

__all__ = ["describe", "check", "run", "load", "Results"]

_PACKAGE_ROOT = str(Path(__file__).resolve().parents[1])
_STATE_KEYS = ("name", "J", "D", "K1", "Ks", "K2", "H", "H_theta", "H_phi")


def describe() -> dict:
    """The models, crystals, coordinate systems, state parameters, units and every default setting."""
    return dict(models=list(MODELS), crystals=sorted(cr.KNOWN), coordinates={k: list(v) for k, v in SIM.COORDS.items()},
                state_parameters=list(_STATE_KEYS), units=["reduced", "meV"], defaults=SIM.default_config())


def _merge(config: dict | None, settings: dict) -> dict:
    cfg = dict(config or {})
    cfg.update(settings)
    return cfg


def _simulation_config(full: dict) -> dict:
    keys = ("model", "crystal", "box", "periodic", "coords", "shape", "states", "units", "T_max", "T_min", "T_step",
            "thermalization", "measurement", "measure_every", "replicas", "config_every", "seed")
    cfg = {k: full[k] for k in keys}
    states = []
    for i, s in enumerate(cfg["states"]):
        s = dict(s)
        s.setdefault("name", f"state-{i + 1}")
        s["J"] = {str(k): float(v) for k, v in (s.get("J") or {1: 1.0}).items()}
        states.append(s)
    cfg["states"] = states
    return cfg


def check(config: dict | None = None, **settings) -> dict:
    """Validate a configuration without running it: {"ok", "errors", "warnings", "config"}; never raises for a
    bad configuration (the errors come back)."""
    cfg = {**SIM.default_config(), **_merge(config, settings)}
    errors, warnings = [], []
    unknown = set(cfg) - set(SIM.default_config()) - {"output_dir", "gpus"}
    if unknown:
        errors.append(f"unknown settings {sorted(unknown)}")
    try:
        full = _simulation_config(cfg)
    except Exception as e:          # noqa: BLE001 - every failure becomes an error message
        return dict(ok=False, errors=errors + [str(e)], warnings=warnings, config=cfg)
    try:
        temps = SIM.temperature_ladder(full["T_max"], full["T_min"], full["T_step"])
        if len(temps) > 1000:
            warnings.append(f"{len(temps)} temperatures")
    except ValueError as e:
        errors.append(str(e))
    try:
        SIM._kB(full["units"])
    except ValueError as e:
        errors.append(str(e))
    for k in ("thermalization", "measurement", "measure_every", "replicas", "config_every"):
        if int(full[k]) < (0 if k == "thermalization" else 1):
            errors.append(f"{k} must be a positive integer")
    names = [s["name"] for s in full["states"]]
    if not names:
        errors.append("define at least one state")
    if len(set(names)) != len(names):
        errors.append("state names must be distinct")
    mask = None
    try:
        _, mask = SIM.active_sites(full["crystal"], full["box"], full["coords"], full["shape"])
    except (ValueError, KeyError) as e:
        errors.append(str(e))
    if mask is not None:
        for i, th in enumerate(full["states"]):
            try:
                s = build_system(full["model"], full["crystal"], full["box"], full["periodic"], th, mask,
                                 log=lambda m: warnings.append(m.strip()))
                if i == 0:
                    cfg["n_sites"], cfg["colouring"] = int(mask.sum()), s.colouring.describe()
            except (ValueError, KeyError) as e:
                errors.append(f"state {i}: {e}")
    return dict(ok=not errors, errors=errors, warnings=warnings, config={**cfg, **full})


def _devices(gpus) -> list[int]:
    import torch
    if not torch.cuda.is_available():
        return []
    usable = list(range(torch.cuda.device_count()))
    if gpus not in (None, "all"):
        wanted = [gpus] if isinstance(gpus, int) else list(gpus)
        usable = [g for g in usable if g in wanted]
    return usable


def run(config: dict | None = None, *, output_dir=None, gpus=None, verbose: bool = True, **settings) -> "Results":
    """Run a simulation on every usable GPU and return its results; refused before anything starts when the
    configuration has errors."""
    cfg = _merge(config, settings)
    if output_dir is not None:
        cfg["output_dir"] = str(output_dir)
    if gpus is not None:
        cfg["gpus"] = gpus
    r = check(cfg)
    if not r["ok"]:
        raise ValueError("the configuration has errors: " + "; ".join(r["errors"]))
    full = r["config"]
    say = print if verbose else (lambda *a, **k: None)
    for w in r["warnings"]:
        say("warning:", w)
    usable = _devices(full.get("gpus", "all"))
    if not usable:
        raise RuntimeError("no HIP device is visible: the MI355X backend has no CPU path")
    out = Path(full.get("output_dir") or "simulation_output")
    sim_cfg = _simulation_config(full)
    prepared = SIM.prepare(sim_cfg, out, log=say)
    say(f"{len(prepared['temperatures'])} temperatures from {prepared['temperatures'][0]} to "
        f"{prepared['temperatures'][-1]}; {len(usable)} GPU(s): {usable}")
    t0 = time.perf_counter()
    n = len(sim_cfg["states"])
    if len(usable) == 1:
        SIM.run(prepared, list(range(n)), out, device=f"cuda:{usable[0]}", log=say)
    else:
        _run_on_gpus(out, n, int(full["replicas"]), usable, say)
    say(f"total: {(time.perf_counter() - t0) / 60:.2f} min")
    return load(out)


def _run_on_gpus(out: Path, n_states: int, replicas: int, usable: list[int], say) -> None:
    """One process per GPU, each with an equal block of every state's replicas (the reference's layout)."""
    blocks = SIM.replica_blocks(replicas, len(usable))
    every = ",".join(map(str, range(n_states)))
    procs = []
    for part, (g, (r0, r1)) in enumerate(zip(usable, blocks)):
        env = dict(os.environ, HIP_VISIBLE_DEVICES=str(g),
                   PYTHONPATH=os.pathsep.join(p for p in (_PACKAGE_ROOT, os.environ.get("PYTHONPATH", "")) if p))
        cmd = [sys.executable, "-u", "-m", "sck_amd.simulate", "--run", str(out), "--states", every,
               "--replicas", f"{r0}:{r1}", "--part", str(part), "--device", "cuda:0"]
        procs.append((g, subprocess.Popen(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        say(f"GPU {g}: replicas {r0}-{r1 - 1} of every state")

    def pump(g, p):
        for line in p.stdout:
            say(f"GPU {g} {line}", end="", flush=True)
    threads = [threading.Thread(target=pump, args=gp) for gp in procs]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    bad = [g for g, p in procs if p.wait() != 0]
    if bad:
        raise RuntimeError(f"the run failed on GPU(s) {bad}: read the lines printed above")


def load(output_dir="simulation_output") -> "Results":
    """The results of a finished run, read from its output folder (the per-GPU files are combined once)."""
    out = Path(output_dir)
    data = SIM.collect(out)
    return Results(out, json.loads((out / "run.json").read_text()), data)


class Results:
    """A finished run: per state and temperature, E and |M| per site, Cv and chi, and the output folder."""

    def __init__(self, output_dir: Path, config: dict, data: dict):
        self.output_dir, self.config, self.arrays = Path(output_dir), config, data
        self.names = [str(n) for n in data["state_names"]]
        self.temperatures = data["temperatures"]
        self.energy_per_site = data["E_site"]
        self.magnetization_per_site = data["M_site"]
        self.specific_heat = data["Cv"]
        self.susceptibility = data["chi"]

    @property
    def temperature_unit(self) -> str:
        return "K" if str(self.config.get("units", "")).lower() == "mev" else "J/k_B"

    def configurations(self):
        """The saved spins, memory-mapped: (states, replicas, saved T, sites, components)."""
        return np.load(self.output_dir / "configurations.npy", mmap_mode="r")

    def table(self) -> list[dict]:
        """One row per state and temperature."""
        return [dict(state=n, T=float(t), E_site=float(self.energy_per_site[i, k]),
                     M_site=float(self.magnetization_per_site[i, k]), Cv=float(self.specific_heat[i, k]),
                     chi=float(self.susceptibility[i, k]))
                for i, n in enumerate(self.names) for k, t in enumerate(self.temperatures)]

    def __repr__(self) -> str:
        T = self.temperatures
        rows = [f"Results({self.output_dir}): {len(self.names)} state(s), {len(T)} temperatures "
                f"{T[0]:g} to {T[-1]:g} {self.temperature_unit}"]
        for i, n in enumerate(self.names):
            rows.append(f"  {n}: at T = {T[-1]:g}  E/site {self.energy_per_site[i, -1]:.5f}  "
                        f"|M| {self.magnetization_per_site[i, -1]:.4f}  Cv {self.specific_heat[i, -1]:.4f}  "
                        f"chi {self.susceptibility[i, -1]:.4f}")
        return "\n".join(rows)
