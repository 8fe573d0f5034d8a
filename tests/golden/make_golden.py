"""Writes the golden fixtures of tests/golden/*.npz by running the reference package's own CPU path.

Run in the development container only (it needs the reference checkout and Triton's CPU interpreter):

    TRITON_INTERPRET=1 python tests/golden/make_golden.py /root/reference/src

Each fixture holds the inputs (system recipe, batch, temperatures, seeds, initial spins) and the outputs of the
reference's Sampler (parallel coloured sweep, which its own tests prove equal to its sequential chain; for a few
systems the per-site sequential chain as well), and float64 energies and magnetizations. Nothing here is run on
the GPU machine and no reference code is copied: the fixtures are data only. The recipes are re-created on the
build's side by tests/golden/recipes.py.
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


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
