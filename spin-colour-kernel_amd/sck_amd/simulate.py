"""Run driver and on-disk format of a simulation (the counterpart of SCK/simulate.py:1-548).

A run is one model, one crystal, one box, one set of active sites (a shape in Cartesian, spherical or
cylindrical coordinates) and a list of states, each a parameter set of the Hamiltonian built by
`sck_amd.hamiltonian.build_system`. Every replica of a state follows one temperature ladder from T_max down to
T_min; at each temperature it thermalizes, then measures every `measure_every` sweeps.

The files are the reference's, so its notebooks and analysis read these runs:
    run.json             the configuration and the derived ladder, sites and units
    sites.npy            (n_sites, 3) Cartesian positions of the active sites
    mask.npy             (b, L3, L2, L1) active sites
    state_XXX_pYY.npz    per state and replica block: r0, r1, M (B, nT, 3), E, E2, Mabs, M2 (B, nT),
                         config_temperatures
    configurations.npy   (n_states, B, n_saved, n_sites, 3) float32 spins every config_every T steps
    magnetization.npy    (n_states, B, nT, 3), written by `collect`
    results.npz          every combined array, Cv and chi included, written by `collect`

On the MI355X a temperature costs a few replays of two captured HIP graphs (a block of thermalization
sweeps; `measure_every` sweeps followed by the float64 energy and magnetization kernels), with the
accumulators about a per-temperature reference energy updated on the device between replays.

Run as a module to process a block of replicas of some states of a prepared run on one device (one process
per GPU, each with its own block of every state's replicas):
    python -m sck_amd.simulate --run DIR --states 0,1 --replicas 0:16 --part 0 --device cuda:0
"""

from __future__ import annotations

import argparse
import ast
import json
import operator
import time
from pathlib import Path

import numpy as np

from . import crystals as cr
from .hamiltonian import build_system

__all__ = ["active_sites", "temperature_ladder", "size_estimate_gb", "simulate_state", "default_config",
           "prepare", "replica_blocks", "run", "collect", "main", "KB_MEV", "COORDS"]

KB_MEV = 0.086173404          # meV / K
COORDS = {"cartesian": ("x", "y", "z"), "spherical": ("r", "theta", "phi"), "cylindrical": ("rho", "phi", "z")}
_MATH = {k: getattr(np, k) for k in ("sqrt", "sin", "cos", "tan", "arctan2", "arccos", "exp", "log", "minimum",
                                     "maximum", "pi", "mod", "floor", "ceil")}
_MATH["abs"] = np.abs
# what a shape may hold: numbers, the coordinate names, the functions above, arithmetic, comparisons and & | ^ ~
_SHAPE_NODES = (ast.Expression, ast.BinOp, ast.UnaryOp, ast.Compare, ast.Call, ast.Name, ast.Load, ast.Constant,
                ast.Add, ast.Sub, ast.Mult, ast.Div, ast.FloorDiv, ast.Mod, ast.Pow, ast.BitAnd, ast.BitOr,
                ast.BitXor, ast.Invert, ast.USub, ast.UAdd, ast.Lt, ast.LtE, ast.Gt, ast.GtE, ast.Eq, ast.NotEq)


def _check_shape(shape: str, names) -> None:
    """Refuse anything in a shape but arithmetic, comparisons, & | ^ ~ and the functions of _MATH
    (SCK/simulate.py:74-95)."""
    try:
        tree = ast.parse(shape, mode="eval")
    except SyntaxError as e:
        raise ValueError(f"the shape {shape!r} is not a valid condition: {e.msg}") from None
    funcs = ", ".join(k for k in _MATH if k != "pi")
    for node in ast.walk(tree):
        if isinstance(node, (ast.BoolOp, ast.Not)):
            raise ValueError("use & (and), | (or), ~ (not) in a shape, with parentheses around each comparison")
        if not isinstance(node, _SHAPE_NODES):
            raise ValueError(f"the shape {shape!r} may hold only numbers, coordinates, operators and {funcs}; "
                             f"found {type(node).__name__}")
        if isinstance(node, ast.Constant) and not isinstance(node.value, (int, float)):
            raise ValueError(f"the shape {shape!r} may hold only numbers as constants")
        if isinstance(node, ast.Call) and not (isinstance(node.func, ast.Name) and node.func.id in _MATH
                                               and not node.keywords):
            raise ValueError(f"the shape {shape!r} calls something other than {funcs}")
        if isinstance(node, ast.Name) and node.id not in _MATH and node.id not in names:
            raise ValueError(f"name {node.id!r} is not defined; the coordinates are {', '.join(names)}")


def _eval_shape(node, env):
    """Evaluate a checked shape tree with Python's operators over numpy arrays, without eval (only the node
    types _check_shape admits reach here)."""
    if isinstance(node, ast.Expression):
        return _eval_shape(node.body, env)
    if isinstance(node, ast.Constant):
        return node.value
    if isinstance(node, ast.Name):
        return env[node.id]
    if isinstance(node, ast.UnaryOp):
        return _UNARY[type(node.op)](_eval_shape(node.operand, env))
    if isinstance(node, ast.BinOp):
        return _BINARY[type(node.op)](_eval_shape(node.left, env), _eval_shape(node.right, env))
    if isinstance(node, ast.Compare):
        left = _eval_shape(node.left, env)
        out = None
        for op, comp in zip(node.ops, node.comparators):
            right = _eval_shape(comp, env)
            r = _COMPARE[type(op)](left, right)
            out = r if out is None else out & r
            left = right
        return out
    if isinstance(node, ast.Call):
        return env[node.func.id](*[_eval_shape(a, env) for a in node.args])
    raise ValueError(f"unsupported shape element {type(node).__name__}")


_UNARY = {ast.USub: operator.neg, ast.UAdd: operator.pos, ast.Invert: operator.invert}
_BINARY = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
           ast.FloorDiv: operator.floordiv, ast.Mod: operator.mod, ast.Pow: operator.pow, ast.BitAnd: operator.and_,
           ast.BitOr: operator.or_, ast.BitXor: operator.xor}
_COMPARE = {ast.Lt: operator.lt, ast.LtE: operator.le, ast.Gt: operator.gt, ast.GtE: operator.ge,
            ast.Eq: operator.eq, ast.NotEq: operator.ne}


def _positions(crystal: cr.Crystal, box) -> np.ndarray:
    """Cartesian position of every site, (b, L3, L2, L1, 3), centred on the box (SCK/simulate.py:99-110)."""
    D = crystal.D
    L1, L2 = int(box[0]), int(box[1])
    L3 = int(box[2]) if D == 3 else 1
    z, y, x = np.meshgrid(np.arange(L3), np.arange(L2), np.arange(L1), indexing="ij")
    cells = np.stack([x, y, z][:D], axis=-1).astype(float)
    out = np.zeros((crystal.b, L3, L2, L1, 3))
    for a in range(crystal.b):
        out[a, ..., :D] = (cells + crystal.basis[a]) @ crystal.vectors
    return out - out.reshape(-1, 3).mean(axis=0)


def active_sites(crystal, box, coords: str = "", shape: str = ""):
    """(positions (b, L3, L2, L1, 3), mask (b, L3, L2, L1)) of a shape in a coordinate system
    (SCK/simulate.py:113-144): coords "cartesian" (x, y, z), "spherical" (r, theta, phi) or "cylindrical"
    (rho, phi, z) about the box centre, angles in radians; shape a condition such as "r <= 10" or
    "(rho <= 8) & (abs(z) <= 2)". Both empty: every site is active."""
    c = crystal if isinstance(crystal, cr.Crystal) else cr.make(str(crystal))
    if len(box) != c.D:
        raise ValueError(f"crystal {c.name} needs a box of {c.D} sizes, got {list(box)}")
    pos = _positions(c, box)
    coords, shape = (coords or "").strip().lower(), (shape or "").strip()
    if not shape:
        if coords and coords not in COORDS:
            raise ValueError(f"unknown coordinate system {coords!r}; known: {', '.join(COORDS)}")
        return pos, np.ones(pos.shape[:-1], bool)
    if not coords:
        raise ValueError("a shape needs its coordinate system: set coords to one of " + ", ".join(COORDS))
    if coords not in COORDS:
        raise ValueError(f"unknown coordinate system {coords!r}; known: {', '.join(COORDS)}")
    x, y, z = pos[..., 0], pos[..., 1], pos[..., 2]
    r = np.sqrt(x * x + y * y + z * z)
    names = {"cartesian": dict(x=x, y=y, z=z),
             "spherical": dict(r=r, theta=np.arccos(np.divide(z, r, out=np.ones_like(z), where=r > 0)),
                               phi=np.arctan2(y, x)),
             "cylindrical": dict(rho=np.hypot(x, y), phi=np.arctan2(y, x), z=z)}[coords]
    _check_shape(shape, names)
    mask = _eval_shape(ast.parse(shape, mode="eval"), {**_MATH, **names})
    mask = np.broadcast_to(np.asarray(mask, bool), pos.shape[:-1]).copy()
    if not mask.any():
        raise ValueError(f"the shape {shape!r} contains no site of the box")
    return pos, mask


def temperature_ladder(T_max: float, T_min: float, T_step: float) -> np.ndarray:
    """Descending temperatures T_max, T_max - T_step, ..., T_min (SCK/simulate.py:243-249)."""
    T_max, T_min, T_step = float(T_max), float(T_min), float(T_step)
    if T_step <= 0 or T_min <= 0 or T_max < T_min:
        raise ValueError("need T_max >= T_min > 0 and T_step > 0")
    k = int(np.floor((T_max - T_min) / T_step + 1e-9))
    return np.round(T_max - T_step * np.arange(k + 1), 10)


def size_estimate_gb(n_states: int, replicas: int, n_saved: int, n_sites: int) -> float:
    """Size of configurations.npy (float32, three components)."""
    return n_states * replicas * n_saved * n_sites * 3 * 4 / 1e9


def _spins(A, S, mask: np.ndarray) -> np.ndarray:
    """(B, n_active, 3) float32 spins of the active sites, missing components zero."""
    x = A.unpack(S).float().permute(0, 2, 3, 4, 5, 1).cpu().numpy()        # (B, b, L3, L2, L1, n)
    x = x[:, mask]
    out = np.zeros(x.shape[:2] + (3,), np.float32)
    out[..., :x.shape[-1]] = x
    return out


def simulate_state(s, replicas: int, temps, n_therm: int, n_meas: int, every: int, config_every: int, seed: int,
                   device="cuda", kB: float = 1.0, log=print) -> dict:
    """The ladder of one state (SCK/simulate.py:322-402): energies are totals, M the mean magnetization per
    active site. Thermalization runs as replays of a graph of 10 k sweeps; each measurement is one replay of
    a graph of k sweeps plus the float64 observables, folded into the accumulators about the temperature's
    reference energy on the device."""
    import torch

    from .sampler import Sampler
    A = Sampler(s, int(replicas), device, seed)
    S = A.init_spins(seed=seed)
    mask = s.site_mask()
    n_act = float(s.n_active)
    k = max(1, int(every))
    m = max(1, int(n_meas) // k)
    big = 10 * k
    ce = max(1, int(config_every))
    temps = np.asarray(temps, float)
    nT, B = len(temps), int(replicas)
    saved = [t for t in range(nT) if t % ce == 0]
    out = dict(M=np.zeros((B, nT, 3), np.float32), E=np.zeros((B, nT)), E2=np.zeros((B, nT)),
               Mabs=np.zeros((B, nT)), M2=np.zeros((B, nT)),
               configs=np.zeros((B, len(saved), int(mask.sum()), 3), np.float32),
               config_temperatures=temps[saved])
    dev = A.device
    f64 = dict(dtype=torch.float64, device=dev)
    E_ref = torch.zeros(B, **f64)
    acc = torch.zeros(4, B, **f64)                 # dE, dE^2, |M|, |M|^2 about the reference
    accv = torch.zeros(B, s.n, **f64)
    A.set_temperature(kB * float(temps[0]))
    rep_big = A.graph(S, big)
    rep_meas = A.graph(S, k, observe=True)

    def fold():
        E = A._obs[:B]
        Mv = A._obs[B:].reshape(B, s.b, s.n).sum(dim=1) / n_act
        d = E - E_ref
        a = Mv.norm(dim=1)
        acc.add_(torch.stack([d, d * d, a, a * a]))
        accv.add_(Mv)

    t0 = time.perf_counter()
    for t, T in enumerate(temps):
        A.set_temperature(kB * float(T))
        for _ in range(int(n_therm) // big):
            rep_big()
        if int(n_therm) % big:
            A.sweep(S, int(n_therm) % big)
        E_ref.copy_(A.observe(S)[0])
        acc.zero_()
        accv.zero_()
        for _ in range(m):
            rep_meas()
            fold()
        dE, dE2, Ma, M2 = (acc / m).cpu().numpy()
        Er = E_ref.cpu().numpy()
        out["E"][:, t] = Er + dE
        out["E2"][:, t] = dE2 + 2 * Er * dE + Er * Er
        out["Mabs"][:, t] = Ma
        out["M2"][:, t] = M2
        out["M"][:, t, :s.n] = (accv / m).float().cpu().numpy()
        if t in saved:
            out["configs"][:, saved.index(t)] = _spins(A, S, mask)
        if t % max(1, nT // 10) == 0 or t == nT - 1:
            log(f"    T = {T:9.4f}   <E>/site = {out['E'][:, t].mean() / n_act:+.4f}"
                f"   <|M|> = {out['Mabs'][:, t].mean():.4f}   {time.perf_counter() - t0:8.1f} s")
    return out


def default_config() -> dict:
    """Every setting of a run with working defaults (a small sc Heisenberg ferromagnet; SCK/simulate.py:405-420)."""
    return dict(model="heisenberg", crystal="sc", box=[12, 12, 12], periodic=[True, True, True], coords="",
                shape="", states=[{"name": "J=1", "J": {"1": 1.0}}], units="reduced", T_max=2.0, T_min=0.2,
                T_step=0.1, thermalization=1000, measurement=2000, measure_every=10, replicas=32, config_every=1,
                seed=1)


def _kB(units: str) -> float:
    u = str(units).strip().lower()
    if u not in ("reduced", "mev"):
        raise ValueError("units must be 'reduced' (kB = 1) or 'meV' (T in kelvin)")
    return 1.0 if u == "reduced" else KB_MEV


def prepare(cfg: dict, out, log=print) -> dict:
    """Validate a configuration, write run.json, sites.npy and mask.npy, and allocate configurations.npy
    (SCK/simulate.py:430-458)."""
    out = Path(out)
    out.mkdir(parents=True, exist_ok=True)
    run_ = {**default_config(), **cfg}
    if not run_["states"]:
        raise ValueError("define at least one state")
    kB = _kB(run_["units"])
    temps = temperature_ladder(run_["T_max"], run_["T_min"], run_["T_step"])
    pos, mask = active_sites(run_["crystal"], run_["box"], run_["coords"], run_["shape"])
    s0 = build_system(run_["model"], run_["crystal"], run_["box"], run_["periodic"], run_["states"][0], mask, log)
    ce = max(1, int(run_["config_every"]))
    n_saved = len(range(0, len(temps), ce))
    n_sites = int(mask.sum())
    run_.update(kB=kB, temperatures=temps.tolist(), periodic_used=list(s0.periodic), n_sites=n_sites,
                n_saved=n_saved, config_temperatures=temps[::ce].tolist())
    for stale in out.glob("state_*.npz"):          # a new run never mixes with an earlier one's blocks
        stale.unlink()
    (out / "run.json").write_text(json.dumps(run_, indent=1))
    np.save(out / "sites.npy", pos[mask].astype(np.float32))
    np.save(out / "mask.npy", mask)
    shape = (len(run_["states"]), int(run_["replicas"]), n_saved, n_sites, 3)
    np.lib.format.open_memmap(out / "configurations.npy", mode="w+", dtype=np.float32, shape=shape).flush()
    log(f"{len(run_['states'])} state(s) x {run_['replicas']} replicas x {len(temps)} temperatures, "
        f"{n_sites} active sites; configurations.npy = {size_estimate_gb(*shape[:4]):.3f} GB")
    return run_


def replica_blocks(replicas: int, parts: int) -> list[tuple[int, int]]:
    """Replicas 0..R-1 in at most `parts` contiguous blocks whose sizes differ by at most one."""
    R, g = int(replicas), max(1, min(int(parts), int(replicas)))
    q, r = divmod(R, g)
    edges = [0]
    for k in range(g):
        edges.append(edges[-1] + q + (1 if k < r else 0))
    return list(zip(edges[:-1], edges[1:]))


def run(cfg: dict, state_indices, out, device="cuda", log=print, replicas=None, part: int = 0) -> None:
    """Simulate states of a prepared run for the replica block [r0, r1) (all replicas by default); the block's
    Philox seed is seed + 1000 state + 100000 part, as the reference's (SCK/simulate.py:471-495). Writes
    state_XXX_pYY.npz and the block's slice of configurations.npy."""
    out = Path(out)
    r0, r1 = (0, int(cfg["replicas"])) if replicas is None else (int(replicas[0]), int(replicas[1]))
    if not 0 <= r0 < r1 <= int(cfg["replicas"]):
        raise ValueError(f"replica block [{r0}, {r1}) is outside 0..{cfg['replicas']}")
    mask = np.load(out / "mask.npy")
    cf = np.load(out / "configurations.npy", mmap_mode="r+")
    for i in state_indices:
        theta = cfg["states"][i]
        log(f"state {i} ({theta.get('name', i)}), replicas {r0}-{r1 - 1}, on {device}")
        s = build_system(cfg["model"], cfg["crystal"], cfg["box"], cfg["periodic_used"], theta, mask, log)
        r = simulate_state(s, r1 - r0, cfg["temperatures"], int(cfg["thermalization"]), int(cfg["measurement"]),
                           int(cfg["measure_every"]), int(cfg["config_every"]),
                           int(cfg["seed"]) + 1000 * int(i) + 100000 * int(part), device, float(cfg["kB"]), log)
        cf[i, r0:r1] = r.pop("configs")
        cf.flush()
        np.savez(out / f"state_{i:03d}_p{int(part):02d}.npz", r0=r0, r1=r1, **r)
    del cf


def _state_parts(out: Path, i: int, R: int) -> list:
    """The replica blocks of state i in replica order; they must cover 0..R-1 exactly once."""
    parts = sorted((np.load(f) for f in out.glob(f"state_{i:03d}_p*.npz")), key=lambda z: int(z["r0"]))
    edge = 0
    for z in parts:
        if int(z["r0"]) != edge:
            break
        edge = int(z["r1"])
    if edge != R or not parts:
        raise ValueError(f"state {i}: its replica blocks do not cover replicas 0..{R - 1} exactly once")
    return parts


def collect(out) -> dict:
    """Combine the state files (SCK/simulate.py:511-530): arrays (n_states, replicas, nT[, 3]) and Cv, chi per
    site (n_states, nT) from the per-replica variances averaged over replicas."""
    out = Path(out)
    cfg = json.loads((out / "run.json").read_text())
    T = np.asarray(cfg["temperatures"])
    N, kB = float(cfg["n_sites"]), float(cfg["kB"])
    R = int(cfg["replicas"])
    states = [_state_parts(out, i, R) for i in range(len(cfg["states"]))]
    res = {k: np.stack([np.concatenate([z[k] for z in parts]) for parts in states])
           for k in ("M", "E", "E2", "Mabs", "M2")}
    varE = np.maximum(res["E2"] - res["E"] ** 2, 0.0)
    varM = np.maximum(res["M2"] - res["Mabs"] ** 2, 0.0)
    combined = dict(temperatures=T, magnetization=res["M"], energy=res["E"], energy2=res["E2"], Mabs=res["Mabs"],
                    M2=res["M2"], Cv=(varE / (kB * T ** 2 * N)).mean(axis=1), chi=(N * varM / (kB * T)).mean(axis=1),
                    E_site=(res["E"] / N).mean(axis=1), M_site=res["Mabs"].mean(axis=1),
                    state_names=np.array([str(s.get("name", i)) for i, s in enumerate(cfg["states"])]))
    np.save(out / "magnetization.npy", combined["magnetization"])
    np.savez(out / "results.npz", **combined)
    return combined


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--run", required=True, help="run directory holding run.json")
    ap.add_argument("--states", required=True, help="comma-separated state indices")
    ap.add_argument("--replicas", default=None, help="replica block r0:r1 (default: every replica)")
    ap.add_argument("--part", type=int, default=0, help="index of the replica block; offsets its seed")
    ap.add_argument("--device", default="cuda:0")
    a = ap.parse_args(argv)
    cfg = json.loads((Path(a.run) / "run.json").read_text())
    block = None if a.replicas is None else tuple(int(x) for x in a.replicas.split(":"))
    run(cfg, [int(x) for x in a.states.split(",") if x != ""], a.run, a.device, lambda m: print(m, flush=True),
        replicas=block, part=a.part)


if __name__ == "__main__":
    main()
