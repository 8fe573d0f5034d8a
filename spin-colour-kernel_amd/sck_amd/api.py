"""The short interface of a simulation: describe, check and run it, then read its results (the reference's
`sck.describe/check/run/load/Results`, SCK/api.py:36-226).

    from sck_amd import api as sck
    sck.describe()                                # models, crystals, state parameters, defaults
    sck.check(crystal="sc", box=[16, 16, 16])     # errors and warnings, nothing runs
    res = sck.run(crystal="sc", box=[16, 16, 16], states=[dict(name="ferromagnet", J={1: 1.0})],
                  T_max=2.0, T_min=0.5, T_step=0.1)
    res.table()                                   # E, |M|, Cv and chi per state and temperature

A configuration is a dict and/or keyword settings (keywords win); settings left out take the defaults of
`describe()["defaults"]`. `run` uses the visible MI355X GPUs: one runs in this process, several run one process
each (`python -m sck_amd.simulate`) on an equal block of every state's replicas. There is no CPU path. The output
folder is the reference's format (sck_amd.simulate).
"""

from __future__ import annotations

import os
import selectors
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

from . import crystals as cr
from . import simulate as SIM
from .hamiltonian import MODELS, build_system, surface

__all__ = ["describe", "check", "run", "load", "Results"]

_PKG_PARENT = str(Path(__file__).resolve().parents[1])
_STATE_KEYS = ("name", "J", "D", "K1", "Ks", "K2", "H", "H_theta", "H_phi")
_RUN_KEYS = tuple(SIM.default_config())            # the settings a run folder records
_EXTRA_KEYS = ("output_dir", "gpus")               # settings of `run` itself


# the energy the state parameters set, term by term (the same model as the reference's, SCK/agent.py:27-30)
_ENERGY = ("E = - sum over shells s of J_s sum over bonds <ij> of shell s of S_i . S_j"
           " - sum over in-plane first-shell bonds of S_i . A_ij S_j, with A_ij the bulk DMI matrix D [u_ij]x of the"
           " unit bond vector u_ij"
           " + sum_i K1_i (sx^2 sy^2 + sy^2 sz^2 + sz^2 sx^2), with K1_i = Ks on surface sites and K1 elsewhere"
           " + K2 sum_i (sx sy sz)^2 - sum_i H . S_i")
_SHAPE_RULES = ("a condition on the coordinates of the chosen system, measured from the centre of the box, angles"
                " in radians: numbers, the coordinate names, pi and the listed functions, arithmetic, comparisons,"
                " and & | ^ ~ with parentheses around each comparison, e.g. (rho <= 10) & (abs(z) <= 2)")


def describe() -> dict:
    """The models, crystals, coordinate systems, shape language, state parameters, units, the energy they
    define and every default setting (the keys of SCK/agent.py:32-46 that concern a run)."""
    return {"hamiltonian": _ENERGY, "models": list(MODELS), "crystals": sorted(cr.KNOWN),
            "coordinates": {k: list(v) for k, v in SIM.COORDS.items()},
            "shape_functions": SIM.shape_functions(), "shape_rules": _SHAPE_RULES,
            "state_parameters": list(_STATE_KEYS), "units": ["reduced", "meV"], "defaults": SIM.default_config()}


def _settings(config, extra) -> dict:
    """Defaults, then the configuration dict, then keyword settings."""
    return {**SIM.default_config(), **dict(config or {}), **extra}


def _normalised(cfg: dict) -> dict:
    """The run's own settings, with every state named ("state-<k>" by default) and J keyed by shell strings."""
    sim = {k: cfg[k] for k in _RUN_KEYS}
    states = []
    for k, st in enumerate(sim["states"], start=1):
        st = dict(st)
        st.setdefault("name", f"state-{k}")
        st["J"] = {str(shell): float(v) for shell, v in (st.get("J") or {1: 1.0}).items()}
        states.append(st)
    sim["states"] = states
    return sim


class _Issues:
    def __init__(self):
        self.errors: list[str] = []
        self.warnings: list[str] = []

    def attempt(self, fn, prefix: str = ""):
        """fn() with its ValueError/KeyError recorded as an error; None then."""
        try:
            return fn()
        except (ValueError, KeyError, TypeError) as e:
            self.errors.append(f"{prefix}{e}")
            return None


def _rule_ladder(sim, issues):
    temps = issues.attempt(lambda: SIM.temperature_ladder(sim["T_max"], sim["T_min"], sim["T_step"]))
    if temps is not None and len(temps) > 1000:
        issues.warnings.append(f"{len(temps)} temperatures")
    issues.attempt(lambda: SIM._kB(sim["units"]))


def _rule_counts(sim, issues):
    for key, least in (("thermalization", 0), ("measurement", 1), ("measure_every", 1), ("replicas", 1),
                       ("config_every", 1)):
        ok = issues.attempt(lambda: int(sim[key]) >= least)
        if ok is False:
            issues.errors.append(f"{key} must be a positive integer")


def _rule_names(sim, issues):
    names = [st["name"] for st in sim["states"]]
    if not names:
        issues.errors.append("define at least one state")
    elif len(set(names)) < len(names):
        issues.errors.append("state names must be distinct")


def _rule_states(sim, issues, report):
    found = issues.attempt(lambda: SIM.active_sites(sim["crystal"], sim["box"], sim["coords"], sim["shape"]))
    if found is None:
        return None
    mask = found[1]
    built = []
    for k, theta in enumerate(sim["states"]):
        s = issues.attempt(lambda: build_system(sim["model"], sim["crystal"], sim["box"], sim["periodic"], theta,
                                                mask, log=lambda m: issues.warnings.append(m.strip())),
                           prefix=f"state {k}: ")
        if s is not None:
            built.append((theta["name"], s))
            if k == 0:
                report.update(n_sites=int(mask.sum()), colouring=s.colouring.describe())
    return mask, built


def _summary(sim, mask, built) -> dict:
    """What a valid configuration amounts to (the report of SCK/agent.py:117-128): its sites (cells of the box,
    active, on the surface, in the bulk), each state's colouring and the periodicity it runs with, the
    temperature ladder, and the size of the run."""
    c = cr.make(str(sim["crystal"]))
    box = [int(v) for v in sim["box"]]
    cells = c.b * int(np.prod(box))
    active = int(mask.sum())
    on_surface = int(surface(c, box, [bool(v) for v in sim["periodic"]], mask).sum())
    temps = SIM.temperature_ladder(sim["T_max"], sim["T_min"], sim["T_step"])
    chains = len(sim["states"]) * int(sim["replicas"])
    saved = len(range(0, len(temps), max(1, int(sim["config_every"]))))
    sweeps = int(sim["thermalization"]) + int(sim["measurement"])
    return {"sites": {"box": cells, "active": active, "surface": on_surface, "bulk": active - on_surface},
            "states": [{"name": name, "colours": s.colouring.n_colours, "colouring": s.colouring.describe(),
                        "periodic_used": list(s.periodic)} for name, s in built],
            "temperatures": {"count": len(temps), "first": float(temps[0]), "last": float(temps[-1])},
            "budget": {"chains": chains, "spin_updates": chains * len(temps) * active * sweeps,
                       "configurations_saved_per_chain": saved,
                       "output_gb": round(SIM.size_estimate_gb(len(sim["states"]), int(sim["replicas"]), saved,
                                                              active), 4)}}


def check(config: dict | None = None, **settings) -> dict:
    """Validate a configuration without running it: {"ok", "errors", "warnings", "config"}, and for a valid one
    also "sites", "states", "temperatures" and "budget". A bad configuration comes back as errors; it never
    raises."""
    cfg = _settings(config, settings)
    issues = _Issues()
    unknown = sorted(set(cfg) - set(_RUN_KEYS) - set(_EXTRA_KEYS))
    if unknown:
        issues.errors.append(f"unknown settings {unknown}")
    sim = issues.attempt(lambda: _normalised(cfg))
    if sim is None:
        return dict(ok=False, errors=issues.errors, warnings=issues.warnings, config=cfg)
    report = {}
    for rule in (_rule_ladder, _rule_counts, _rule_names):
        rule(sim, issues)
    found = _rule_states(sim, issues, report)
    out = dict(ok=not issues.errors, errors=issues.errors, warnings=issues.warnings,
               config={**cfg, **report, **sim})
    if out["ok"]:
        out.update(_summary(sim, *found))
    return out


def _visible_gpus(wanted) -> list[int]:
    import torch
    if not torch.cuda.is_available():
        return []
    ids = range(torch.cuda.device_count())
    if wanted in (None, "all"):
        return list(ids)
    keep = {wanted} if isinstance(wanted, int) else set(wanted)
    return [g for g in ids if g in keep]


def run(config: dict | None = None, *, output_dir=None, gpus=None, verbose: bool = True, **settings) -> "Results":
    """Check a configuration, run it on the usable GPUs and return its results. Nothing starts when the check
    reports errors (ValueError); without a HIP device it raises RuntimeError."""
    extra = dict(settings)
    if output_dir is not None:
        extra["output_dir"] = str(output_dir)
    if gpus is not None:
        extra["gpus"] = gpus
    report = check(config, **extra)
    if not report["ok"]:
        raise ValueError("the configuration has errors: " + "; ".join(report["errors"]))
    cfg = report["config"]
    say = print if verbose else (lambda *a, **k: None)
    for w in report["warnings"]:
        say("warning:", w)
    devices = _visible_gpus(cfg.get("gpus", "all"))
    if not devices:
        raise RuntimeError("no HIP device is visible: the MI355X backend has no CPU path")
    out = Path(cfg.get("output_dir") or "simulation_output")
    prepared = SIM.prepare(_normalised(cfg), out, log=say)
    T = prepared["temperatures"]
    say(f"{len(T)} temperatures from {T[0]} to {T[-1]}; {len(devices)} GPU(s): {devices}")
    start = time.perf_counter()
    states = list(range(len(prepared["states"])))
    if len(devices) == 1:
        SIM.run(prepared, states, out, device=f"cuda:{devices[0]}", log=say)
    else:
        _one_process_per_gpu(out, states, int(prepared["replicas"]), devices, say)
    say(f"total: {(time.perf_counter() - start) / 60:.2f} min")
    return load(out)


def _one_process_per_gpu(out: Path, states, replicas: int, devices: list[int], say) -> None:
    """One `python -m sck_amd.simulate` process per GPU, each on its replica block of every state, started with
    only that GPU visible; their output lines are relayed as they come, prefixed by the GPU."""
    path = os.pathsep.join(p for p in (_PKG_PARENT, os.environ.get("PYTHONPATH", "")) if p)
    sel = selectors.DefaultSelector()
    procs = {}
    for part, (g, (r0, r1)) in enumerate(zip(devices, SIM.replica_blocks(replicas, len(devices)))):
        cmd = [sys.executable, "-u", "-m", "sck_amd.simulate", "--run", str(out),
               "--states", ",".join(map(str, states)), "--replicas", f"{r0}:{r1}", "--part", str(part),
               "--device", "cuda:0"]
        p = subprocess.Popen(cmd, env={**os.environ, "HIP_VISIBLE_DEVICES": str(g), "PYTHONPATH": path},
                             stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        procs[g] = p
        sel.register(p.stdout, selectors.EVENT_READ, g)
        say(f"GPU {g}: replicas {r0}-{r1 - 1} of every state")
    open_streams = len(procs)
    while open_streams:
        for key, _ in sel.select():
            line = key.fileobj.readline()
            if line:
                say(f"GPU {key.data} {line}", end="", flush=True)
            else:
                sel.unregister(key.fileobj)
                open_streams -= 1
    failed = [g for g, p in procs.items() if p.wait() != 0]
    if failed:
        raise RuntimeError(f"the run failed on GPU(s) {failed}: read the lines printed above")


def load(output_dir="simulation_output") -> "Results":
    """The results of a finished run, read from its folder (the per-GPU files are combined on the way)."""
    folder = SIM.RunFolder(output_dir)
    return Results(folder.path, folder.config(), SIM.collect(folder.path))


class Results:
    """A finished run. Per state (rows) and temperature (columns): energy_per_site, magnetization_per_site,
    specific_heat, susceptibility. Also names, temperatures, config (run.json), arrays (results.npz) and
    output_dir; configurations() maps the saved spins without reading them."""

    def __init__(self, output_dir: Path, config: dict, data: dict):
        self.output_dir = Path(output_dir)
        self.config = config
        self.arrays = data
        self.names = [str(v) for v in data["state_names"]]
        self.temperatures = data["temperatures"]

    energy_per_site = property(lambda self: self.arrays["E_site"])
    magnetization_per_site = property(lambda self: self.arrays["M_site"])
    specific_heat = property(lambda self: self.arrays["Cv"])
    susceptibility = property(lambda self: self.arrays["chi"])

    @property
    def temperature_unit(self) -> str:
        return "K" if str(self.config.get("units", "")).strip().lower() == "mev" else "J/k_B"

    def configurations(self):
        """The saved spins, memory-mapped: (states, replicas, saved temperatures, sites, 3)."""
        return SIM.RunFolder(self.output_dir).configurations()

    def table(self) -> list[dict]:
        """One row per (state, temperature)."""
        cols = (("E_site", self.energy_per_site), ("M_site", self.magnetization_per_site),
                ("Cv", self.specific_heat), ("chi", self.susceptibility))
        return [{"state": name, "T": float(T), **{k: float(v[i, j]) for k, v in cols}}
                for i, name in enumerate(self.names) for j, T in enumerate(self.temperatures)]

    def plot(self, path=None, show: bool = True):
        """Energy, |M|, Cv and chi per site against temperature, one line per state, in a 2 x 2 figure; saved
        to `path` when one is given (needs matplotlib)."""
        import matplotlib.pyplot as plt
        panels = (("energy per site", self.energy_per_site), ("|M| per site", self.magnetization_per_site),
                  ("Cv per site", self.specific_heat), ("chi per site", self.susceptibility))
        fig, axes = plt.subplots(2, 2, figsize=(11, 7), sharex=True)
        for ax, (label, values) in zip(axes.flat, panels):
            for name, row in zip(self.names, values):
                ax.plot(self.temperatures, row, marker="o", ms=3, lw=1.2, label=name)
            ax.set_ylabel(label)
            ax.grid(alpha=0.25)
        for ax in axes[1]:
            ax.set_xlabel(f"T ({self.temperature_unit})")
        axes[0, 0].legend(fontsize=9)
        c = self.config
        fig.suptitle(f"{c.get('model')} on {c.get('crystal')} {list(c.get('box', []))}: {len(self.names)} state(s), "
                     f"{c.get('replicas')} replicas")
        fig.tight_layout()
        if path is not None:
            fig.savefig(path, dpi=150)
        if show:
            plt.show()
        return fig

    def __repr__(self) -> str:
        T, u = self.temperatures, self.temperature_unit
        head = f"Results({self.output_dir}): {len(self.names)} state(s), {len(T)} temperatures {T[0]:g} to {T[-1]:g} {u}"
        last = [f"  {name}: T = {T[-1]:g}  E/site {self.energy_per_site[i, -1]:.5f}  "
                f"|M| {self.magnetization_per_site[i, -1]:.4f}  Cv {self.specific_heat[i, -1]:.4f}  "
                f"chi {self.susceptibility[i, -1]:.4f}" for i, name in enumerate(self.names)]
        return "\n".join([head] + last)
