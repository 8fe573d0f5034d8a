"""Run driver and on-disk format of a simulation: the caller of the sweep path (the counterpart of
SCK/simulate.py:1-548).

A run is one model, one crystal, one box, one set of active sites (a condition in Cartesian, spherical or
cylindrical coordinates about the box centre) and a list of states, each a parameter set of the Hamiltonian
(`sck_amd.hamiltonian.build_system`). Every replica of a state walks one temperature ladder from T_max down to
T_min; at each temperature it thermalizes, then observes every `measure_every` sweeps.

Files of a run folder. Their names, array keys and shapes are the reference's, so its notebooks and analysis
read MI355X runs:
    run.json             the settings plus the derived ladder, units and site count
    sites.npy            (n_sites, 3) Cartesian positions of the active sites
    mask.npy             (b, L3, L2, L1) active sites
    state_XXX_pYY.npz    state XXX, replica block YY: r0, r1, M (B, nT, 3), E, E2, Mabs, M2 (B, nT),
                         config_temperatures
    configurations.npy   (n_states, B, n_saved, n_sites, 3) float32 spins, every config_every temperatures
    magnetization.npy    (n_states, B, nT, 3), from `collect`
    results.npz          the combined arrays with Cv and chi, from `collect`

On the GPU a temperature costs a few replays of two captured HIP graphs, one of thermalization sweeps and one of
`measure_every` sweeps followed by the float64 energy and magnetization reductions; the accumulators, taken about
a reference energy per temperature, stay on the device until the temperature is done.

A module run processes one replica block of some states of a prepared run on one device (one process per GPU):
    python -m sck_amd.simulate --run DIR --states 0,1 --replicas 0:16 --part 0 --device cuda:0
"""

from __future__ import annotations

import argparse
import ast
import json
import math
import operator
import time
from dataclasses import asdict, dataclass, field
from pathlib import Path

import numpy as np

from . import crystals as cr
from .hamiltonian import build_system

__all__ = ["active_sites", "site_positions", "temperature_ladder", "size_estimate_gb", "simulate_state",
           "default_config", "prepare", "replica_blocks", "run", "collect", "main", "RunFolder", "KB_MEV",
           "COORDS"]

KB_MEV = 0.086173404          # Boltzmann's constant in meV / K: temperatures in kelvin when units = "meV"
COORDS = {"cartesian": ("x", "y", "z"), "spherical": ("r", "theta", "phi"), "cylindrical": ("rho", "phi", "z")}


# ----------------------------------------------------------------------------------------------- shapes
_FUNCS = {"sqrt": np.sqrt, "sin": np.sin, "cos": np.cos, "tan": np.tan, "arctan2": np.arctan2,
          "arccos": np.arccos, "exp": np.exp, "log": np.log, "minimum": np.minimum, "maximum": np.maximum,
          "mod": np.mod, "floor": np.floor, "ceil": np.ceil, "abs": np.abs}
_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
        ast.FloorDiv: operator.floordiv, ast.Mod: operator.mod, ast.Pow: operator.pow, ast.BitAnd: operator.and_,
        ast.BitOr: operator.or_, ast.BitXor: operator.xor, ast.USub: operator.neg, ast.UAdd: operator.pos,
        ast.Invert: operator.invert, ast.Lt: operator.lt, ast.LtE: operator.le, ast.Gt: operator.gt,
        ast.GtE: operator.ge, ast.Eq: operator.eq, ast.NotEq: operator.ne}


def shape_functions() -> list[str]:
    """The names a shape condition may call, and the constant pi."""
    return list(_FUNCS) + ["pi"]


class _Shape(ast.NodeVisitor):
    """A shape condition, checked and evaluated in one walk of its syntax tree: numbers, the coordinate names,
    the functions of _FUNCS (and pi), arithmetic, comparisons and & | ^ ~ are allowed, anything else is refused
    with the reference's messages (SCK/simulate.py:74-95). No `eval`: each node is computed with numpy."""

    def __init__(self, text: str, coords: dict):
        self.text, self.coords = text, coords
        self.funcs = ", ".join(_FUNCS)

    def evaluate(self):
        try:
            tree = ast.parse(self.text, mode="eval")
        except SyntaxError as e:
            raise ValueError(f"the shape {self.text!r} is not a valid condition: {e.msg}") from None
        return self.visit(tree.body)

    def generic_visit(self, node):
        raise ValueError(f"the shape {self.text!r} may hold only numbers, coordinates, operators and {self.funcs}; "
                         f"found {type(node).__name__}")

    def visit_BoolOp(self, node):
        raise ValueError("use & (and), | (or), ~ (not) in a shape, with parentheses around each comparison")

    def visit_UnaryOp(self, node):
        if isinstance(node.op, ast.Not):
            return self.visit_BoolOp(node)
        return self._op(node.op)(self.visit(node.operand))

    def visit_BinOp(self, node):
        return self._op(node.op)(self.visit(node.left), self.visit(node.right))

    def visit_Compare(self, node):
        result, left = True, self.visit(node.left)
        for op, rhs in zip(node.ops, node.comparators):
            right = self.visit(rhs)
            result = result & self._op(op)(left, right)
            left = right
        return result

    def visit_Constant(self, node):
        if isinstance(node.value, bool) or not isinstance(node.value, (int, float)):
            raise ValueError(f"the shape {self.text!r} may hold only numbers as constants")
        return node.value

    def visit_Name(self, node):
        if node.id in self.coords:
            return self.coords[node.id]
        if node.id == "pi":
            return np.pi
        if node.id in _FUNCS:
            raise ValueError(f"the shape {self.text!r} uses the function {node.id} without calling it")
        raise ValueError(f"name {node.id!r} is not defined; the coordinates are {', '.join(self.coords)}")

    def visit_Call(self, node):
        if not isinstance(node.func, ast.Name) or node.func.id not in _FUNCS or node.keywords:
            raise ValueError(f"the shape {self.text!r} calls something other than {self.funcs}")
        return _FUNCS[node.func.id](*[self.visit(a) for a in node.args])

    def _op(self, op):
        f = _OPS.get(type(op))
        if f is None:
            self.generic_visit(op)
        return f


def site_positions(crystal: cr.Crystal, box) -> np.ndarray:
    """Cartesian position of every site, (b, L3, L2, L1, 3), with the origin at the mean of all sites."""
    D = crystal.D
    dims = [int(v) for v in box] + [1] * (3 - D)                  # (L1, L2, L3)
    grid = np.indices(dims[::-1]).transpose(1, 2, 3, 0)[..., ::-1][..., :D]     # (L3, L2, L1, D): x, y[, z]
    frac = grid[None].astype(float) + np.asarray(crystal.basis, float)[:, None, None, None, :]
    pos = np.zeros(frac.shape[:-1] + (3,))
    pos[..., :D] = frac @ np.asarray(crystal.vectors, float)
    return pos - pos.reshape(-1, 3).mean(axis=0)


def _coordinate_arrays(system: str, pos: np.ndarray) -> dict:
    x, y, z = np.moveaxis(pos, -1, 0)
    if system == "cartesian":
        return dict(x=x, y=y, z=z)
    if system == "cylindrical":
        return dict(rho=np.hypot(x, y), phi=np.arctan2(y, x), z=z)
    r = np.sqrt(x * x + y * y + z * z)
    cos_t = np.divide(z, r, out=np.ones_like(z), where=r > 0)
    return dict(r=r, theta=np.arccos(cos_t), phi=np.arctan2(y, x))


def active_sites(crystal, box, coords: str = "", shape: str = ""):
    """(positions (b, L3, L2, L1, 3), mask (b, L3, L2, L1)) of the sites where `shape` holds
    (SCK/simulate.py:113-144). coords: "cartesian" (x, y, z), "spherical" (r, theta, phi) or "cylindrical"
    (rho, phi, z), about the box centre, angles in radians; shape: a condition such as "r <= 10" or
    "(rho <= 8) & (abs(z) <= 2)". With no shape every site is active."""
    c = crystal if isinstance(crystal, cr.Crystal) else cr.make(str(crystal))
    if len(box) != c.D:
        raise ValueError(f"crystal {c.name} needs a box of {c.D} sizes, got {list(box)}")
    pos = site_positions(c, box)
    system, text = (coords or "").strip().lower(), (shape or "").strip()
    if system and system not in COORDS:
        raise ValueError(f"unknown coordinate system {system!r}; known: {', '.join(COORDS)}")
    if not text:
        return pos, np.ones(pos.shape[:-1], bool)
    if not system:
        raise ValueError("a shape needs its coordinate system: set coords to one of " + ", ".join(COORDS))
    cond = _Shape(text, _coordinate_arrays(system, pos)).evaluate()
    mask = np.array(np.broadcast_to(np.asarray(cond, bool), pos.shape[:-1]))
    if not mask.any():
        raise ValueError(f"the shape {text!r} contains no site of the box")
    return pos, mask


# ------------------------------------------------------------------------------------------ the ladder
def temperature_ladder(T_max: float, T_min: float, T_step: float) -> np.ndarray:
    """T_max, T_max - T_step, ... down to T_min (inclusive), each rounded to 10 decimals."""
    hi, lo, dt = float(T_max), float(T_min), float(T_step)
    if not (dt > 0 and lo > 0 and hi >= lo):
        raise ValueError("need T_max >= T_min > 0 and T_step > 0")
    steps = math.floor((hi - lo) / dt + 1e-9)
    return np.round(hi - dt * np.arange(steps + 1), 10)


def size_estimate_gb(n_states: int, replicas: int, n_saved: int, n_sites: int) -> float:
    """GB of configurations.npy: float32 spins of three components."""
    return 12.0 * n_states * replicas * n_saved * n_sites / 1e9


class _Ladder:
    """One state's temperature ladder on the device: a HIP graph of `10 k` thermalization sweeps, a HIP graph of
    k sweeps followed by the float64 observables, and accumulators of (E - E_ref, (E - E_ref)^2, |M|, |M|^2) and
    of the magnetization vector, with E_ref the energy at the start of each temperature's measurements (the
    reference's accumulators, SCK/simulate.py:349-398)."""

    def __init__(self, system, replicas: int, seed: int, device, every: int, kB: float):
        import torch

        from .sampler import Sampler
        self.torch = torch
        self.s, self.kB = system, float(kB)
        self.A = Sampler(system, int(replicas), device, seed)
        self.S = self.A.init_spins(seed=seed)
        self.k = max(1, int(every))
        self.N = float(system.n_active)
        f64 = dict(dtype=torch.float64, device=self.A.device)
        B, n = self.A.B, system.n
        self.E_ref = torch.zeros(B, **f64)
        self.sums = torch.zeros(4, B, **f64)
        self.mvec = torch.zeros(B, n, **f64)
        self.thermal_block = None
        self.measure_step = None

    def _graphs(self) -> None:
        if self.thermal_block is None:
            self.thermal_block = self.A.graph(self.S, 10 * self.k)
            self.measure_step = self.A.graph(self.S, self.k, observe=True)

    def thermalize(self, sweeps: int) -> None:
        block = 10 * self.k
        for _ in range(sweeps // block):
            self.thermal_block()
        if sweeps % block:
            self.A.sweep(self.S, sweeps % block)

    def measure(self, count: int) -> dict:
        """`count` observations, one per k sweeps; the per-replica means of the accumulated quantities."""
        A, B, s = self.A, self.A.B, self.s
        self.E_ref.copy_(A.observe(self.S)[0])
        self.sums.zero_()
        self.mvec.zero_()
        E, Mraw = A.observed()
        for _ in range(count):
            self.measure_step()
            m = Mraw.sum(dim=1) / self.N                           # (B, n): magnetization per active site
            dE = E - self.E_ref
            a = m.norm(dim=1)
            self.sums.add_(self.torch.stack((dE, dE * dE, a, a * a)))
            self.mvec.add_(m)
        mean = (self.sums / count).cpu().numpy()
        Er = self.E_ref.cpu().numpy()
        return dict(E=Er + mean[0], E2=mean[1] + 2.0 * Er * mean[0] + Er * Er, Mabs=mean[2], M2=mean[3],
                    M=(self.mvec / count).float().cpu().numpy())

    def snapshot(self, mask: np.ndarray) -> np.ndarray:
        """(B, n_active, 3) float32 spins of the active sites, missing components zero."""
        x = self.A.unpack(self.S).float().movedim(1, -1).cpu().numpy()[:, mask]     # (B, n_active, n)
        return np.pad(x, ((0, 0), (0, 0), (0, 3 - x.shape[-1])))

    def run(self, temps, n_therm: int, n_meas: int, config_every: int, log) -> dict:
        temps = np.asarray(temps, float)
        nT, B = len(temps), self.A.B
        mask = self.s.site_mask()
        saved = list(range(0, nT, max(1, int(config_every))))
        out = dict(M=np.zeros((B, nT, 3), np.float32), E=np.zeros((B, nT)), E2=np.zeros((B, nT)),
                   Mabs=np.zeros((B, nT)), M2=np.zeros((B, nT)),
                   configs=np.zeros((B, len(saved), int(mask.sum()), 3), np.float32),
                   config_temperatures=temps[saved])
        count = max(1, int(n_meas) // self.k)
        t0 = time.perf_counter()
        for t, T in enumerate(temps):
            self.A.set_temperature(self.kB * float(T))
            self._graphs()
            self.thermalize(int(n_therm))
            r = self.measure(count)
            for key in ("E", "E2", "Mabs", "M2"):
                out[key][:, t] = r[key]
            out["M"][:, t, :self.s.n] = r["M"]
            if t in saved:
                out["configs"][:, saved.index(t)] = self.snapshot(mask)
            if t == nT - 1 or t % max(1, nT // 10) == 0:
                log(f"    T = {T:9.4f}   <E>/site = {out['E'][:, t].mean() / self.N:+.4f}"
                    f"   <|M|> = {out['Mabs'][:, t].mean():.4f}   {time.perf_counter() - t0:8.1f} s")
        return out


def simulate_state(s, replicas: int, temps, n_therm: int, n_meas: int, every: int, config_every: int, seed: int,
                   device="cuda", kB: float = 1.0, log=print) -> dict:
    """The ladder of one state (SCK/simulate.py:322-402): per replica and temperature the mean total energy E,
    E^2, |M| and |M|^2 per site and the magnetization vector M per site, and the saved configurations."""
    return _Ladder(s, replicas, seed, device, every, kB).run(temps, n_therm, n_meas, config_every, log)


# ------------------------------------------------------------------------------------------- run folder
@dataclass
class _Settings:
    """Every setting of a run and its default (the reference's, SCK/simulate.py:405-420): a small periodic sc
    Heisenberg ferromagnet."""
    model: str = "heisenberg"
    crystal: str = "sc"
    box: list = field(default_factory=lambda: [12, 12, 12])
    periodic: list = field(default_factory=lambda: [True, True, True])
    coords: str = ""
    shape: str = ""
    states: list = field(default_factory=lambda: [{"name": "J=1", "J": {"1": 1.0}}])
    units: str = "reduced"
    T_max: float = 2.0
    T_min: float = 0.2
    T_step: float = 0.1
    thermalization: int = 1000
    measurement: int = 2000
    measure_every: int = 10
    replicas: int = 32
    config_every: int = 1
    seed: int = 1


def default_config() -> dict:
    """Every setting of a run with its default."""
    return asdict(_Settings())


def _kB(units: str) -> float:
    table = {"reduced": 1.0, "mev": KB_MEV}
    u = str(units).strip().lower()
    if u not in table:
        raise ValueError("units must be 'reduced' (kB = 1) or 'meV' (T in kelvin)")
    return table[u]


class RunFolder:
    """The files of one run (see the module docstring); all reads and writes of the format go through here."""

    def __init__(self, path):
        self.path = Path(path)

    def config(self) -> dict:
        return json.loads((self.path / "run.json").read_text())

    def start(self, cfg: dict, positions: np.ndarray, mask: np.ndarray, configs_shape) -> None:
        self.path.mkdir(parents=True, exist_ok=True)
        for old in self.path.glob("state_*.npz"):         # a new run never mixes with an earlier one's blocks
            old.unlink()
        (self.path / "run.json").write_text(json.dumps(cfg, indent=1))
        np.save(self.path / "sites.npy", positions[mask].astype(np.float32))
        np.save(self.path / "mask.npy", mask)
        np.lib.format.open_memmap(self.path / "configurations.npy", mode="w+", dtype=np.float32,
                                  shape=tuple(configs_shape)).flush()

    def mask(self) -> np.ndarray:
        return np.load(self.path / "mask.npy")

    def configurations(self, writable: bool = False):
        return np.load(self.path / "configurations.npy", mmap_mode="r+" if writable else "r")

    def write_block(self, state: int, part: int, r0: int, r1: int, arrays: dict) -> None:
        np.savez(self.path / f"state_{state:03d}_p{part:02d}.npz", r0=r0, r1=r1, **arrays)

    def blocks(self, state: int, replicas: int) -> list:
        """The replica blocks of a state in replica order, which must tile 0..replicas-1 exactly."""
        files = [np.load(f) for f in self.path.glob(f"state_{state:03d}_p*.npz")]
        files.sort(key=lambda z: int(z["r0"]))
        bounds = [(int(z["r0"]), int(z["r1"])) for z in files]
        starts = [0] + [hi for _, hi in bounds[:-1]]
        if not files or [lo for lo, _ in bounds] != starts or bounds[-1][1] != replicas:
            raise ValueError(f"state {state}: its replica blocks do not cover replicas 0..{replicas - 1} exactly once")
        return files

    def write_results(self, combined: dict) -> None:
        np.save(self.path / "magnetization.npy", combined["magnetization"])
        np.savez(self.path / "results.npz", **combined)


def prepare(cfg: dict, out, log=print) -> dict:
    """Check a configuration, write run.json, sites.npy and mask.npy and allocate configurations.npy
    (SCK/simulate.py:432-458); returns the completed configuration."""
    settings = {**default_config(), **cfg}
    if not settings["states"]:
        raise ValueError("define at least one state")
    kB = _kB(settings["units"])
    temps = temperature_ladder(settings["T_max"], settings["T_min"], settings["T_step"])
    pos, mask = active_sites(settings["crystal"], settings["box"], settings["coords"], settings["shape"])
    first = build_system(settings["model"], settings["crystal"], settings["box"], settings["periodic"],
                         settings["states"][0], mask, log)
    saved = temps[::max(1, int(settings["config_every"]))]
    n_sites = int(mask.sum())
    settings.update(kB=kB, temperatures=temps.tolist(), periodic_used=list(first.periodic), n_sites=n_sites,
                    n_saved=len(saved), config_temperatures=saved.tolist())
    shape = (len(settings["states"]), int(settings["replicas"]), len(saved), n_sites, 3)
    RunFolder(out).start(settings, pos, mask, shape)
    log(f"{shape[0]} state(s) x {shape[1]} replicas x {len(temps)} temperatures, {n_sites} active sites; "
        f"configurations.npy = {size_estimate_gb(*shape[:4]):.3f} GB")
    return settings


def replica_blocks(replicas: int, parts: int) -> list[tuple[int, int]]:
    """Replicas 0..R-1 as at most `parts` contiguous blocks, the first R mod parts of them one larger."""
    R = int(replicas)
    g = max(1, min(int(parts), R))
    sizes = [R // g + (k < R % g) for k in range(g)]
    edges = np.cumsum([0] + sizes).tolist()
    return list(zip(edges[:-1], edges[1:]))


def run(cfg: dict, state_indices, out, device="cuda", log=print, replicas=None, part: int = 0) -> None:
    """Simulate states of a prepared run for the replica block [r0, r1) (all replicas by default). A block's
    Philox base seed is seed + 1000 state + 100000 part, the reference's (SCK/simulate.py:471-495), so blocks
    run on different devices are independent chains. Writes state_XXX_pYY.npz and the block's configurations."""
    folder = RunFolder(out)
    R = int(cfg["replicas"])
    r0, r1 = (0, R) if replicas is None else (int(replicas[0]), int(replicas[1]))
    if not 0 <= r0 < r1 <= R:
        raise ValueError(f"replica block [{r0}, {r1}) is outside 0..{R}")
    mask = folder.mask()
    store = folder.configurations(writable=True)
    for i in state_indices:
        theta = cfg["states"][i]
        log(f"state {i} ({theta.get('name', i)}), replicas {r0}-{r1 - 1}, on {device}")
        s = build_system(cfg["model"], cfg["crystal"], cfg["box"], cfg["periodic_used"], theta, mask, log)
        seed = int(cfg["seed"]) + 1000 * int(i) + 100000 * int(part)
        res = simulate_state(s, r1 - r0, cfg["temperatures"], int(cfg["thermalization"]), int(cfg["measurement"]),
                             int(cfg["measure_every"]), int(cfg["config_every"]), seed, device, float(cfg["kB"]),
                             log)
        store[i, r0:r1] = res.pop("configs")
        store.flush()
        folder.write_block(int(i), int(part), r0, r1, res)
    del store


def collect(out) -> dict:
    """Combine the state files of a run (SCK/simulate.py:511-530): arrays (n_states, replicas, nT[, 3]), and Cv
    and chi per site (n_states, nT) as the replica means of var(E) / (kB T^2 N) and N var(|M|) / (kB T)."""
    folder = RunFolder(out)
    cfg = folder.config()
    T = np.asarray(cfg["temperatures"], float)
    N, kB, R = float(cfg["n_sites"]), float(cfg["kB"]), int(cfg["replicas"])
    per_state = [folder.blocks(i, R) for i in range(len(cfg["states"]))]

    def stacked(key):
        return np.stack([np.concatenate([z[key] for z in blocks]) for blocks in per_state])
    E, E2, Mabs, M2, M = (stacked(k) for k in ("E", "E2", "Mabs", "M2", "M"))
    var_E = np.clip(E2 - E * E, 0.0, None)
    var_M = np.clip(M2 - Mabs * Mabs, 0.0, None)
    combined = dict(temperatures=T, magnetization=M, energy=E, energy2=E2, Mabs=Mabs, M2=M2,
                    Cv=np.mean(var_E / (kB * N * T * T), axis=1), chi=np.mean(N * var_M / (kB * T), axis=1),
                    E_site=np.mean(E / N, axis=1), M_site=np.mean(Mabs, axis=1),
                    state_names=np.array([str(s.get("name", i)) for i, s in enumerate(cfg["states"])]))
    folder.write_results(combined)
    return combined


def main(argv=None):
    p = argparse.ArgumentParser(description="Simulate a replica block of some states of a prepared run.")
    p.add_argument("--run", required=True, help="the run folder (holds run.json)")
    p.add_argument("--states", required=True, help="state indices, comma separated")
    p.add_argument("--replicas", default=None, help="replica block r0:r1; every replica when left out")
    p.add_argument("--part", type=int, default=0, help="block index (offsets the Philox seed)")
    p.add_argument("--device", default="cuda:0")
    a = p.parse_args(argv)
    cfg = RunFolder(a.run).config()
    block = None if a.replicas is None else tuple(int(v) for v in a.replicas.split(":"))
    states = [int(v) for v in a.states.split(",") if v.strip()]
    run(cfg, states, a.run, a.device, lambda m: print(m, flush=True), replicas=block, part=a.part)


if __name__ == "__main__":
    main()
