"""Writes the golden fixtures of tests/golden/*.npz by running the reference package's own CPU path.

Run in the development container only (it needs the reference checkout and Triton's CPU interpreter):

    TRITON_INTERPRET=1 python tests/golden/make_golden.py /root/reference/src

Each trajectory fixture holds the inputs (system recipe, batch, temperatures, seeds, initial spins) and the outputs
of the reference's Sampler (parallel coloured sweep, which its own tests prove equal to its sequential chain; for a few
systems the per-site sequential chain as well), and float64 energies and magnetizations. Nothing here is run on
the GPU machine and no reference code is copied: the fixtures are data only. The recipes are re-created on the
build's side by tests/golden/recipes.py.

The driver fixtures (build_system.npz, run_ising.npz, run_heis.npz, anneal_nanodot.npz) hold what the reference's
Hamiltonian builder, run driver (prepare / run / collect) and anneal ladder write for the inputs in BUILD_CASES,
RUN_CASES and ANNEAL_CASE below.
"""

from __future__ import annotations

import json
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
from recipes import RECIPES  # noqa: E402


def main(ref_src: str) -> None:
    os.environ.setdefault("TRITON_INTERPRET", "1")
    sys.path.insert(0, ref_src)
    import torch
    from Spin_Colour_Kernel import bitwise as bw
    from Spin_Colour_Kernel import crystals as cr
    from Spin_Colour_Kernel import system as sy
    from Spin_Colour_Kernel.sampler import Sampler

    for name, rc in RECIPES.items():
        out = HERE / f"{name}.npz"
        if out.exists():
            continue
        t0 = time.time()
        s = eval(rc["system"], {"sy": sy, "cr": cr, "np": np})
        if rc.get("acceptance"):
            s.acceptance = rc["acceptance"]
        kw = dict(rc.get("sampler", {}))
        A = Sampler(s, rc["B"], device="cpu", seed=rc["seed"], **kw)
        T = torch.as_tensor(rc["T"], dtype=torch.float32)
        A.set_temperature(T)
        S = A.init_spins(seed=rc["init_seed"])
        rec = dict(S0=S.numpy().copy(), beta=A.beta.numpy().copy(), state0=A.state.numpy().copy())
        if A.bitwise:
            rec["ptab"] = A.ptab.numpy().copy()
        mode = rc.get("mode", "sweep")
        steps = []
        for k in range(rc["sweeps"]):
            if mode == "sync":
                A.sync_sweep(S, 1)
            elif mode == "reference":
                A.reference_sweep(S, 1)
            else:
                A.sweep(S, 1)
            steps.append(S.numpy().copy())
        rec["S"] = np.stack(steps)
        if not A.bitwise and not A.packed:
            rec["E"] = A.energy(S).numpy()
            rec["M"] = A.magnetization(S).numpy()
        else:
            x = A.unpack(S)
            rec["unpacked"] = x.numpy()
            rec["E"] = A.energy(S).numpy()
            rec["M"] = A.magnetization(S).numpy()
        rec["meta"] = np.frombuffer(json.dumps(rc).encode(), np.uint8)
        np.savez_compressed(out, **rec)
        print(f"{name:28s} {time.time() - t0:7.1f} s", flush=True)
    # bitwise tables of the reference at a few inverse temperatures (known answers for the table builders)
    out = HERE / "bitwise_tables.npz"
    if not out.exists():
        beta = np.array([0.25, 0.5, 1.0 / 2.269185, 1.0])
        tabs = {}
        for key, make in (("square", lambda: sy.ising("square", 4)),
                          ("sc_h", lambda: sy.System("h", 1, cr.make("sc"), (4, 4, 4), (True,) * 3, {1: 1.0}, h=(0.3,)))):
            for acc in ("metropolis", "barker"):
                s = make()
                s.acceptance = acc
                tabs[f"{key}_{acc}_f"] = bw.p_table(s, beta)
                tabs[f"{key}_{acc}_i"] = bw.p_fixed(s, beta)
        np.savez_compressed(out, beta=beta, **tabs)
    # Philox: the reference's NumPy twin at a few keys and counters (pins the counter/key convention)
    out = HERE / "philox.npz"
    if not out.exists():
        ctr = np.arange(64, dtype=np.int64) * 7 + 3
        np.savez_compressed(out, ctr=ctr, u=bw.philox_uniform4(123457, ctr),
                            raw=bw.philox_raw4(99, ctr, ctr % 5 + 1),
                            seeds=np.array([bw.sweep_seed(b, w) for b in (0, 1, 11, 123456) for w in (0, 1, 999, 10 ** 6)]),
                            lane=bw.lane_uniforms(5, 3, 2, 16), uni=bw.uniforms(5, 3, 6, 16))
    driver_fixtures()


# ------------------------------------------------------------------ the callers of the path: driver fixtures
# build_system states: (model, crystal, box, periodic, theta, shape condition or "") -- written as data so that
# tests/test_simulate.py rebuilds the same inputs
BUILD_CASES = {
    "heis_sc_all_terms": ("heisenberg", "sc", [4, 4, 4], [True, True, True],
                          {"name": "all", "J": {"1": 1.0, "2": 0.2}, "D": 0.3, "K1": 0.1, "Ks": 0.25, "K2": 0.05,
                           "H": 0.4, "H_theta": 30.0, "H_phi": 45.0}, ("", "")),
    "heis_bcc_sphere": ("heisenberg", "bcc", [4, 4, 4], [False, False, False],
                        {"D": 0.2, "K1": 0.15, "Ks": 0.4, "H": 0.3, "H_theta": 90.0, "H_phi": 10.0},
                        ("spherical", "r <= 1.6")),
    "xy_triangular_field": ("xy", "triangular", [6, 6], [True, True], {"J": {"1": -1.0}, "H": 0.5, "H_phi": 60.0},
                            ("", "")),
    "ising_square_odd_fallback": ("ising", "square", [5, 5], [True, True], {"J": {"1": 1.0}, "H": 0.7}, ("", "")),
    "heis_kagome_dmi": ("heisenberg", "kagome", [4, 4], [True, True], {"J": {"1": 1.0}, "D": 0.25, "K1": 0.05},
                        ("cartesian", "x > -2.5")),
    "heis_fcc_ks_only": ("heisenberg", "fcc", [2, 2, 2], [True, True, True], {"Ks": 0.3}, ("", "")),
}

# run folders: a configuration for the reference's prepare / run / collect on its CPU path
RUN_CASES = {
    "run_ising": dict(model="ising", crystal="square", box=[8, 8], periodic=[True, True], coords="", shape="",
                      states=[{"name": "ferro", "J": {"1": 1.0}}, {"name": "field", "J": {"1": 1.0}, "H": 0.3}],
                      units="reduced", T_max=3.0, T_min=2.0, T_step=0.5, thermalization=10, measurement=10,
                      measure_every=5, replicas=4, config_every=2, seed=3),
    "run_heis": dict(model="heisenberg", crystal="sc", box=[4, 4, 4], periodic=[True, True, True],
                     coords="spherical", shape="r <= 1.8",
                     states=[{"name": "dmi", "J": {"1": 1.0}, "D": 0.2, "K1": 0.1, "Ks": 0.2, "H": 0.3,
                              "H_theta": 20.0}],
                     units="meV", T_max=8.0, T_min=6.0, T_step=2.0, thermalization=2, measurement=4,
                     measure_every=2, replicas=2, config_every=1, seed=5),
}

# one anneal ladder: nanodot parameter row [Jex2, Jex3, Jex4, Kan1, KanS, Hex, KDM] and its settings
ANNEAL_CASE = dict(params=[0.0, 0.0, 0.0, 0.177, 0.139, 0.483, 0.88], Rd=2.0, L=2, angle=30.0, replicas=2,
                   temps_K=[12.0, 6.0], n_therm=2, n_meas=4, every=2, seed=9)


def _system_record(s) -> dict:
    bonds = sorted(s.extra)
    return dict(n=np.int64(s.n), box=np.asarray(s.box), periodic=np.asarray(s.periodic),
                J_shells=np.asarray(sorted(s.J)), J_values=np.asarray([s.J[k] for k in sorted(s.J)]),
                extra_bonds=np.asarray(bonds, np.int64),
                extra=np.asarray([s.extra[i] for i in bonds]) if bonds else np.zeros((0, s.n, s.n)),
                h=np.asarray(s.h if s.h is not None else [], float),
                k1=np.asarray(s.cubic_k1 if s.cubic_k1 is not None else np.zeros(0), np.float32),
                k2=np.float64(s.cubic_k2), mask=s.site_mask(), name=np.array(s.name))


def driver_fixtures() -> None:
    """build_system states, two run folders and one anneal ladder, written by the reference's own functions
    (SCK/simulate.py:198-240, 432-530; SCK/anneal.py:40-66, 143-234) on its CPU path."""
    import tempfile

    from Spin_Colour_Kernel import anneal as AN
    from Spin_Colour_Kernel import simulate as SIM
    out = HERE / "build_system.npz"
    if not out.exists():
        rec = {}
        for key, (model, crystal, box, per, theta, (coords, shape)) in BUILD_CASES.items():
            _, mask = SIM.active_sites(crystal, box, coords, shape)
            s = SIM.build_system(model, crystal, box, per, theta, mask, log=lambda m: None)
            rec.update({f"{key}/{k}": v for k, v in _system_record(s).items()})
        np.savez_compressed(out, **rec)
        print(f"{'build_system':28s} written", flush=True)
    for key, cfg in RUN_CASES.items():
        out = HERE / f"{key}.npz"
        if out.exists():
            continue
        t0 = time.time()
        with tempfile.TemporaryDirectory() as d:
            run = SIM.prepare(cfg, d, log=lambda m: None)
            SIM.run(run, list(range(len(cfg["states"]))), d, device="cpu", log=lambda m: None)
            res = SIM.collect(d)
            rec = {"run_json": np.array((Path(d) / "run.json").read_text()),
                   "sites": np.load(Path(d) / "sites.npy"), "mask": np.load(Path(d) / "mask.npy"),
                   "configurations": np.load(Path(d) / "configurations.npy")}
            for i in range(len(cfg["states"])):
                z = np.load(Path(d) / f"state_{i:03d}_p00.npz")
                rec.update({f"state_{i:03d}/{k}": z[k] for k in z.files})
            rec.update({f"results/{k}": v for k, v in res.items()})
        np.savez_compressed(out, **rec)
        print(f"{key:28s} {time.time() - t0:7.1f} s", flush=True)
    out = HERE / "anneal_nanodot.npz"
    if not out.exists():
        t0 = time.time()
        c = ANNEAL_CASE
        s = AN.build_state(c["params"], c["Rd"], c["L"], c["angle"], (1,))
        r = AN.anneal_state(s, c["replicas"], c["temps_K"], c["n_therm"], c["n_meas"], c["every"], c["seed"],
                            device="cpu", log=lambda m: None)
        red = AN.reduce_obs(r["obs"], c["temps_K"], s.n_active)
        np.savez_compressed(out, obs=r["obs"], mvec=r["mvec"], snaps=r["snaps"], n_active=np.int64(s.n_active),
                            **{f"reduced/{k}": v for k, v in red.items()})
        print(f"{'anneal_nanodot':28s} {time.time() - t0:7.1f} s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
