"""Throughput of the multi-colour sweep on one MI355X: site updates/s and fraction of the HBM roofline, and
effective samples/s of the learned proposal width against fixed-width Metropolis.

Default workload (the largest single-GPU configuration of the headline metric): XY model on the triangular
lattice, L = 2048 periodic (4 colour classes), J1 = 1, h = 0, fp32 2-vectors with uniform random angles,
B = 2048 replicas = 16 temperatures x 128 in [0.7, 1.1]. A step is one coloured sweep over all replicas; inputs
are resident in HBM before the timed region, which replays a HIP graph of exactly `--steps` sweeps.

    python bench.py [--steps K] [--warmup W] [--workload xy|bitwise|ising|heisenberg|nanodisks|learned|...]

Prints one JSON line. `roofline` is for the dominant kernel: algorithmic bytes per launch (compulsory read +
write of each spin, 16 B per XY site update) over the launch's average duration, measured with HIP events on
the stream the kernels run on; `valu` gives the same kernel's vector-instruction issue fraction from its
committed counter run. `cpu_baseline` times the CPU oracle (the same chain, sequential, in C) on a bounded
sample on this host. `ess` (default workload only) is the learned-versus-fixed comparison on the Heisenberg
bcc 32^3 configuration (B = 8192): effective samples/s of the energy, learned width over the best of five
fixed widths and over the uniform redraw.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
for p in (str(REPO), str(REPO / "spin-colour-kernel_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (description, builder, sampler kwargs, B, temperatures, algorithmic bytes per update)
    "bitwise": "bitwise Ising sc L=128 periodic, B=1024 (32 words), independent draws, T=4.5",
    "ising": "Ising square L=64 periodic, int8, B=512 (8 temps x 64 in [1.5,3.5]), seed 0",
    "heisenberg": "Heisenberg bcc L=32 J1=1 J2=-0.2 Du=0.1 z h=0.05 z, fp32, B=8192 (32 temps x 256 in [0.5,3.0])",
    "xy": "XY triangular L=2048 periodic J1=1, fp32, B=2048 (16 temps x 128 in [0.7,1.1])",
    "learned": "Heisenberg bcc L=32 J1=1 J2=-0.2 Du=0.1 z h=0.05 z, B=8192 (32 temps in [0.5,3.0]): effective "
               "samples/s of the learned cone width over the best of five fixed widths (by measured ESS/s)",
    "nanodisks": "64 Heisenberg sc nanodisks, radius U[8,40], thickness U{4..12} (numpy default_rng(2024)), open, "
                 "J1=1, bulk DMI D=0.1, Ks=0.3 along the surface normal, 16 replicas each (T in [0.3,1.8]), fp32, "
                 "every disk in one launch per class, graph-captured sweeps, autotuned schedule",
    "learned-xy": "XY triangular L=2048 periodic J1=1, B=512 by default (--batch; 16 temps in [0.7,1.1]; the "
                  "comparison holds up to three copies of the lattice and a tiled workspace, so 2048 replicas "
                  "do not fit), ordered start, measured through the tiled sweep's running totals: effective "
                  "samples/s of the learned width over the best of four fixed widths, tau_int of E and |M| per "
                  "temperature",
    "reference-table": "the reference's own throughput benchmark: its 13 systems at B = 8..128 and its 2-D bitwise "
                       "scan, measured by its protocol, each next to its published H100 rate (one JSON line each)",
}

# Published H100 rates of the reference's throughput benchmark, flips/ns (= site updates per ns):
# results-product/Results/tables/e6-throughput-h100.csv:2-14 (systems x B = 8, 16, 32, 64, 128) and
# results-product/Results/tables/e6-scan.csv:2-5 ("Ours r10": bitwise independent, B = 64, side = 8 L).
REF_H100 = {
    "Ising square": (46.954, 57.2, 62.518, 67.982, 71.55),
    "Ising triangular": (20.283, 24.666, 30.7, 35.838, 38.894),
    "Ising sc": (19.349, 22.336, 50.054, 55.961, 59.688),
    "Ising bcc": (38.318, 44.268, 91.613, 101.497, 108.281),
    "Ising fcc": (39.567, 45.069, 75.4, 81.86, 85.83),
    "XY square": (38.418, 45.008, 49.17, 50.606, 52.334),
    "XY sc": (18.782, 21.942, 23.99, 25.159, 25.626),
    "Heisenberg sc": (18.594, 21.786, 23.921, 24.886, 25.539),
    "Heisenberg bcc": (37.083, 43.334, 46.689, 48.963, 50.691),
    "Heisenberg fcc": (37.715, 42.585, 38.799, 40.888, 41.902),
    "Heisenberg triangular AFM": (14.866, 17.204, 18.879, 19.578, 20.031),
    "Heisenberg kagome AFM": (49.637, 58.17, 58.33, 56.66, 57.862),
    "Heisenberg nanodot": (5.05, 8.14, 10.127, 12.025, 13.433),
}
REF_H100_BATCHES = (8, 16, 32, 64, 128)
REF_H100_SCAN = {512: 572.8, 1024: 698.8, 2048: 745.2, 4096: 761.1}
# single-replica large lattices: the chunked-row check (e1-chunk-exactness.csv:2-5, "Chunked flips/ns": eager
# 3-sweep wall-clock rate, seed 7, T = 1, initial spins seed 8) and the reference kernel's own side of its
# CUDA-METRO pairing (e6-cudametro-v17.csv:2,6,10, float32: Heisenberg triangular ferromagnet, T = 1, seed 6007,
# graph replays of k = ceil(2e7 / N) sweeps)
REF_H100_LARGE = {("heisenberg", "triangular", 2048): 3.993, ("heisenberg", "triangular", 4096): 4.43,
                  ("xy", "square", 2048): 5.3, ("heisenberg", "sc-rows", 1152): 0.41}
REF_H100_METRO = {1024: 11.747, 2048: 12.11, 4096: 12.585}


def _make(workload: str):
    import numpy as np
    import torch

    from sck_amd import crystals as cr
    from sck_amd import system as sy
    if workload == "bitwise":
        s = sy.ising("sc", 128)
        return s, dict(bitwise=True, independent=True), 1024, torch.full((1024,), 4.5), 0.25
    if workload == "ising":
        s = sy.ising("square", 64)
        T = torch.as_tensor(np.repeat(np.linspace(1.5, 3.5, 8).astype(np.float32), 64))
        return s, dict(packed=False), 512, T, 2.0
    if workload == "heisenberg":
        s = sy.System("Heisenberg bcc J1J2", 3, cr.make("bcc", (1, 2)), (32, 32, 32), (True,) * 3,
                      {1: 1.0, 2: -0.2}, h=(0.0, 0.0, 0.05), uniaxial=(0.1, (0.0, 0.0, 1.0)))
        T = torch.as_tensor(np.repeat(np.linspace(0.5, 3.0, 32).astype(np.float32), 256))
        return s, {}, 8192, T, 24.0
    if workload == "xy":
        s = sy.xy("triangular", 2048)
        T = torch.as_tensor(np.repeat(np.linspace(0.7, 1.1, 16).astype(np.float32), 128))
        return s, {}, 2048, T, 16.0
    raise SystemExit(f"unknown workload {workload!r}; known: {', '.join(WORKLOADS)}")


def prebuilt_specs():
    """(system, mode, batch, replica_major, tiled) of every bench workload, for the build step's kernel
    pre-compilation (the layout is the one the Sampler picks for it)."""
    from sck_amd import hipgen as G
    from sck_amd.sampler import RESIDENT_MIN_BATCH, RESIDENT_SMALL_SITES, choose_layout
    out = []
    for wl in ("bitwise", "ising", "heisenberg", "xy"):
        s, kw, B, _, _ = _make(wl)
        if kw.get("bitwise"):
            out.append((s, G.MODE_BITWISE_INDEP, B, False, False))
            continue
        mode = G.MODE_INT8 if s.n == 1 else G.MODE_F32
        res = B >= RESIDENT_MIN_BATCH or s.nsite <= RESIDENT_SMALL_SITES
        layout, tiled = choose_layout(s, mode, B, res)
        out.append((s, mode, B, layout == "replica", tiled))
    return out


def _cpu_model() -> str:
    """The host CPU as /proc/cpuinfo names it (model name, family, model, stepping)."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("model name", "cpu family", "model", "stepping") and k not in info:
                info[k] = v.strip()
    except OSError:
        pass
    name = info.get("model name", "unknown")
    rest = ", ".join(f"{k} {info[k]}" for k in ("cpu family", "model", "stepping") if k in info)
    return f"{name} ({rest})" if rest else name


def _host_cores() -> dict:
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"nproc": os.cpu_count(), "usable": usable}


def cpu_baseline(workload: str, s, A, S0, budget_s: float = 10.0) -> dict:
    """Time the CPU oracle (the sequential chain in C) on a bounded sample of the same workload: one thread
    (the reported value) and, as a second row, one thread per replica (word) on all usable host cores."""
    import numpy as np

    import oracle as O
    os_ = O.OracleSystem(s)
    ncores = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1))

    def run(nrep: int, threads: int, budget: float):
        if A.bitwise:
            x = O.unpack(S0[0:nrep], 32 * nrep, s.site_mask())[:, 0]
            tab = A.ptab.cpu().numpy()[0:nrep]
        else:
            beta = A.beta.cpu().numpy()
            x = S0[list(range(nrep))].astype(np.float32)
        t0 = time.perf_counter()
        sweeps = 0
        while time.perf_counter() - t0 < budget and sweeps < 1000:
            if A.bitwise:
                x = O.bitwise_chain(os_, x, tab, 0, sweeps, 1, independent=A.independent, threads=threads)
            else:
                x, _ = O.chain(os_, x, beta, 0, sweeps, 1, reps=list(range(nrep)), threads=threads)
            sweeps += 1
        dt = time.perf_counter() - t0
        per = 32 if A.bitwise else 1
        return nrep * per * s.n_active * sweeps / dt, sweeps, dt

    unit = "word (32 replicas)" if A.bitwise else "replica"
    ncores = min(ncores, int(S0.shape[0]))
    v1, sw1, dt1 = run(1, 1, budget_s)
    vn, swn, dtn = run(ncores, ncores, budget_s)
    return {"value": v1, "unit": "site updates/s", "cores": 1, "kind": "port",
            "sample": f"1 {unit} x {s.n_active} sites x {sw1} sweeps, sequential chain, {dt1:.1f} s",
            "cpu_model": _cpu_model(), "host_cores": _host_cores(),
            "all_cores": {"value": vn, "cores": ncores,
                          "sample": f"{ncores} x 1 {unit} (one per thread, OpenMP) x {s.n_active} sites x {swn} "
                                    f"sweeps, {dtn:.1f} s"}}


def ordered_spins(A):
    """All spins along the first axis (masked sites zero), written on the device in the sampler's storage. Below
    the BKT temperature an ordered start relaxes by spin waves (the energy within tens of sweeps), where a
    random start spends thousands of sweeps annihilating vortices on a 2048^2 lattice."""
    import torch
    s, n = A.sys, A.sys.n
    m = torch.as_tensor(s.site_mask(), dtype=torch.float32, device=A.device)          # (b, L3, L2, L1)
    if A.replica_major:
        store = torch.zeros((s.b, s.L3, s.L2, s.L1, n, A.B), dtype=torch.float32, device=A.device)
        store[..., 0, :] = m[..., None]
        return store.permute(5, 4, 0, 1, 2, 3)
    S = torch.zeros((A.B, n) + s.shape, dtype=torch.float32, device=A.device)
    S[:, 0] = m
    return S


def device_spins(A, seed: int = 0):
    """Random XY (uniform angle) or Heisenberg (normalised Gaussian) spins drawn on the device one block at a
    time with torch's seeded CUDA generator, straight into the sampler's storage (a full-size batch is too
    large to draw quickly on the host)."""
    import math

    import torch
    s, n = A.sys, A.sys.n
    g = torch.Generator(device=A.device).manual_seed(int(seed))

    def draw(shape_lead):                      # (..., n) unit vectors
        if n == 2:
            phi = torch.rand(shape_lead, generator=g, device=A.device) * (2 * math.pi)
            return torch.stack([torch.cos(phi), torch.sin(phi)], dim=-1)
        x = torch.randn(shape_lead + (n,), generator=g, device=A.device)
        return x / x.norm(dim=-1, keepdim=True)
    if A.replica_major:
        store = torch.empty((s.b, s.L3, s.L2, s.L1, n, A.B), dtype=torch.float32, device=A.device)
        for a in range(s.b):
            for z in range(s.L3):
                for y0 in range(0, s.L2, 64):
                    y1 = min(s.L2, y0 + 64)
                    v = draw((y1 - y0, s.L1, A.B))                       # (rows, L1, B, n)
                    store[a, z, y0:y1] = v.permute(0, 1, 3, 2)
        return store.permute(5, 4, 0, 1, 2, 3)
    S = torch.empty((A.B, n) + s.shape, dtype=torch.float32, device=A.device)
    for r0 in range(0, A.B, 16):
        v = draw((min(16, A.B - r0),) + s.shape)                          # (reps, b, L3, L2, L1, n)
        S[r0:r0 + 16] = v.permute(0, 5, 1, 2, 3, 4)
    return S


def traffic_from_profiles(workload: str):
    """HBM bytes per launch of the class kernel from the committed rocprofv3 counter summary, if any."""
    f = REPO / "profiles" / f"traffic_{workload}.json"
    if f.exists():
        return json.loads(f.read_text()).get("bytes_per_launch")
    return None


def valu_from_profiles(workload: str):
    """Vector-instruction issue fraction of the dominant kernel from its committed rocprofv3 counter run
    (4 x SQ_INSTS_VALU / (1024 SIMDs x cycles per XCD)), if any."""
    f = REPO / "profiles" / f"valu_{workload}.json"
    if f.exists():
        d = json.loads(f.read_text())
        return {"frac": d.get("valu_issue_frac"), "salu_frac": d.get("salu_issue_frac"),
                "valu_per_update": d.get("valu_per_update"), "source": f"profiles/valu_{workload}.json"}
    return None


def _graph_rate(A, S, k: int, upd: float, reps: int) -> float:
    """Median flips/ns over `reps` measurements of >= ~100 ms of replays of a graph of k sweeps, the
    reference's protocol (SCK/exp_perf.py:471-505, _events_ms at :22-30)."""
    import math

    import numpy as np
    import torch
    rp = A.graph(S, k)
    for _ in range(2):
        rp()
    torch.cuda.synchronize()
    ms0 = rp.time_ms(1)
    nrep = max(1, int(math.ceil(100.0 / ms0)))
    return float(np.median([upd * nrep / (rp.time_ms(nrep) * 1e6) for _ in range(reps)]))


def reference_table_main(a) -> None:
    """The reference's throughput benchmark on this backend: the 13 systems of SCK/exp_perf.py:60-74 at
    B = 8, 16, 32, 64, 128 with the Sampler's default options (packed storage for Ising from B = 32, as the
    reference), T = the system's reference temperature, seed 1, graph replays of k = ceil(2e7 / (B N))
    sweeps; then the 2-D bitwise scan of SCK/exp_perf.py:202-231 (Ising square L = side / 8, B = 64,
    independent draws, T = Tc, seed 5011); then the single-replica large lattices of the reference's
    chunked-row check (SCK/chunk_check.py) and of its own kernel's side of the CUDA-METRO pairing
    (SCK/cudametro_pair.py), each with that experiment's protocol. One JSON line per row with the published H100
    rate beside it, then the geometric mean of the ratios."""
    import math

    import numpy as np
    import torch

    from sck_amd import system as sy
    from sck_amd.sampler import Sampler
    systems = [
        sy.ising("square", 256, T_ref=sy.TC["Ising square"]),
        sy.ising("triangular", 192, T_ref=sy.TC["Ising triangular"]),
        sy.ising("sc", 32, T_ref=sy.TC["Ising sc"]),
        sy.ising("bcc", 32, T_ref=sy.TC["Ising bcc"]),
        sy.ising("fcc", 32, T_ref=sy.TC["Ising fcc"]),
        sy.xy("square", 256, T_ref=0.893),
        sy.xy("sc", 32, T_ref=sy.TC["XY sc"]),
        sy.heisenberg("sc", 32, T_ref=sy.TC["Heisenberg sc"]),
        sy.heisenberg("bcc", 32, T_ref=sy.TC["Heisenberg bcc"]),
        sy.heisenberg("fcc", 32, T_ref=1.0),
        sy.heisenberg("triangular", 192, J=-1.0, T_ref=0.285, label="Heisenberg triangular AFM"),
        sy.heisenberg("kagome", 192, J=-1.0, T_ref=0.1, label="Heisenberg kagome AFM"),
        sy.nanodot(),
    ]
    reps = 5
    ratios = []
    for s in systems:
        for B, ref in zip(REF_H100_BATCHES, REF_H100[s.name]):
            A = Sampler(s, B, "cuda:0", 1)
            A.set_temperature(s.T_ref or 1.0)
            S = A.init_spins(seed=1)
            k = max(1, int(math.ceil(2e7 / (B * s.n_active))))
            r = _graph_rate(A, S, k, B * s.n_active * k, reps)
            ratios.append(r / ref)
            print(json.dumps({"benchmark": "reference throughput table", "system": s.name, "batch": B,
                              "active_sites": s.n_active, "storage": A.layout if not A.packed else "packed",
                              "flips_per_ns": r, "h100_flips_per_ns": ref, "ratio": r / ref}), flush=True)
            del A, S
            torch.cuda.empty_cache()
    for L, ref in REF_H100_SCAN.items():
        s = sy.ising("square", L, T_ref=sy.TC["Ising square"])
        A = Sampler(s, 64, "cuda:0", 5011, bitwise=True, independent=True)
        A.set_temperature(sy.TC["Ising square"])
        S = A.init_spins(seed=5011)
        k = max(2, int(2e10 // (64 * L * L)) // 10)
        r = _graph_rate(A, S, k, 64 * L * L * k, reps)
        ratios.append(r / ref)
        print(json.dumps({"benchmark": "reference bitwise scan", "system": f"bitwise Ising square L={L}", "batch": 64,
                          "flips_per_ns": r, "h100_flips_per_ns": ref, "ratio": r / ref}), flush=True)
        del A, S
        torch.cuda.empty_cache()
    from sck_amd import crystals as cr
    for (kind, crystal, L), ref in REF_H100_LARGE.items():
        # SCK/chunk_check.py:21-50: B = 1, four sweeps, then the wall-clock rate of three eager sweeps
        if crystal == "sc-rows":
            s = sy.System("Heisenberg sc long rows", 3, cr.make("sc"), (L, 8, 8), (True,) * 3, {1: 1.0})
        else:
            s = (sy.heisenberg if kind == "heisenberg" else sy.xy)(crystal, L)
        A = Sampler(s, 1, "cuda:0", 7)
        A.set_temperature(1.0)
        S = A.init_spins(seed=8)
        A.sweep(S, 4)
        rs = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            A.sweep(S, 3)
            torch.cuda.synchronize()
            rs.append(s.n_active * 3 / ((time.perf_counter() - t0) * 1e9))
        r = float(np.median(rs))
        ratios.append(r / ref)
        print(json.dumps({"benchmark": "reference large-lattice rows", "system": s.name, "box": list(s.box), "batch": 1,
                          "kernel": "tiled" if A.use_tiled else A.layout, "flips_per_ns": r,
                          "h100_flips_per_ns": ref, "ratio": r / ref}), flush=True)
        del A, S
        torch.cuda.empty_cache()
    for L, ref in REF_H100_METRO.items():
        # SCK/cudametro_pair.py:368-400: the ferromagnet, B = 1 in float32, T = 1, seed 6007
        s = sy.heisenberg("triangular", L)
        A = Sampler(s, 1, "cuda:0", 6007)
        A.set_temperature(1.0)
        S = A.init_spins(seed=6007)
        k = max(1, int(math.ceil(2e7 / s.n_active)))
        r = _graph_rate(A, S, k, s.n_active * k, reps)
        ratios.append(r / ref)
        print(json.dumps({"benchmark": "reference kernel of the CUDA-METRO pairing", "system": s.name, "box": list(s.box),
                          "batch": 1, "kernel": "tiled" if A.use_tiled else A.layout, "flips_per_ns": r,
                          "h100_flips_per_ns": ref, "ratio": r / ref}), flush=True)
        del A, S
        torch.cuda.empty_cache()
    print(json.dumps({"benchmark": "reference throughput table, bitwise scan and large lattices", "rows": len(ratios),
                      "geomean_ratio_vs_h100": float(np.exp(np.mean(np.log(ratios)))),
                      "min_ratio": float(min(ratios)), "max_ratio": float(max(ratios))}), flush=True)


def learned_main(a) -> None:
    """Effective independent samples/s of E and |M|, learned width vs the best fixed width (one JSON line)."""
    import numpy as np

    from sck_amd import system as sy
    from sck_amd.ess import compare_proposals
    widths, fine = (0.5, 1.0, 1.5, 2.0, 2.5), True
    if a.workload == "learned-xy":
        s, B = sy.xy("triangular", 2048), a.batch or 512
        temps = np.linspace(0.7, 1.1, 16).astype(np.float32)
        equil = 300
        widths, fine = (1.0, 1.5, 2.0, 2.5), False      # the widths that bracket the best; no table study
    else:
        s, _, B, T, _ = _make("heisenberg")
        temps = np.unique(T.numpy())
        equil = 1000
    meas = max(50, a.steps)
    print(f"comparing proposals on {s.name}, B={B}", flush=True)
    r = compare_proposals(s, B, temps, warmup=200, equil=equil, therm=max(20, a.warmup), n_meas=meas,
                          log=lambda m: print(m, flush=True), widths=widths, fine_table=fine,
                          init=ordered_spins if a.workload == "learned-xy" else None)
    out = {"metric": "effective samples/s, learned vs fixed proposal", "value": r["gain_E"],
           "unit": "ratio of ESS/s of the energy, geometric mean over the temperature ladder", "n_gpus": 1, "steps": meas, "warmup": a.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (random initial spins from the seeded torch CPU generator)",
           "config": {"workload": WORKLOADS[a.workload], "batch": B, "sites": s.n_active},
           "gain_E": r["gain_E"], "gain_M": r["gain_M"], "gain_E_bottleneck": r["gain_E_bottleneck"],
           "gain_M_bottleneck": r["gain_M_bottleneck"], "gain_E_sum": r["gain_E_sum"], "gain_M_sum": r["gain_M_sum"],
           "fixed_width": r["fixed_width"], "fixed_scan_gm_ess_per_s": r["fixed_scan_gm_ess"],
           "vs_uniform_redraw": {k: r["vs_uniform"][k] for k in ("gain_E", "gain_M", "gain_E_bottleneck",
                                                                   "gain_M_bottleneck")},
           "ess_per_s": {k: {"E": r[k]["ess_E_per_s"], "M": r[k]["ess_M_per_s"],
                             "updates_per_s": r[k]["updates_per_s"], "fused": r[k]["fused"]}
                         for k in ("uniform", "fixed", "learned")},
           "per_T_best_fixed_over_best_single_fixed": r["gain_E_per_T_best_over_fixed"],
           "gain_E_vs_per_T_best_fixed": r["gain_E_vs_per_T_best"],
           "observables_max_z": r["max_z"], "observables_frac_within_2sigma": r["frac_within_2sigma"],
           "temperatures": [float(t) for t in temps],
           "tau_int": {k: {"E": [float(r[k]["tau_E"][j]) for j in range(len(temps))],
                           "M": [float(r[k]["tau_M"][j]) for j in range(len(temps))]}
                       for k in ("uniform", "fixed", "learned")},
           "train_seconds": r["train_seconds"]}
    if a.out_dir:
        f = Path(a.out_dir)
        f.mkdir(parents=True, exist_ok=True)
        (f / f"{a.workload.replace('-', '_')}_comparison.json").write_text(json.dumps(r, default=float, indent=1))
    print(json.dumps(out, default=float), flush=True)


def nanodisk_main(a) -> None:
    """Batched uneven systems: all 64 nanodisks x 16 replicas swept together (one JSON line)."""
    import numpy as np
    import torch

    import oracle as O
    from sck_amd.batch import BatchSampler
    from sck_amd.hamiltonian import nanodisk_set
    systems = nanodisk_set(64, 8.0, 40.0, 4, 12, seed=2024)
    reps = 16
    sch = "auto"
    if a.schedule != "auto":
        from sck_amd.batch import Schedule
        kind, block, tile, order, grid = a.schedule.split("/")
        sch = Schedule(kind, int(block), int(tile), order, int(grid))
    Bt = BatchSampler(systems, reps, "cuda:0", seed=0, schedule=sch)
    temps = np.linspace(0.3, 1.8, reps).astype(np.float32)
    Bt.set_temperature([temps] * len(systems))
    S = Bt.init_spins(seed=0)
    S0 = [v.cpu().numpy() for v in Bt.views(S)]
    if sch == "auto":
        Bt.autotune(S, log=lambda m: print(m, file=sys.stderr, flush=True))
    k = 10
    rep = Bt.graph(S, k)
    for _ in range(max(1, a.warmup // k)):
        rep()
    torch.cuda.synchronize()
    nrep = max(1, a.steps // k)
    steps = nrep * k
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(nrep):
        rep()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sec = e0.elapsed_time(e1) / 1e3
    upd = reps * sum(s.n_active for s in systems)
    value = upd * steps / sec
    sch = Bt.schedule
    # dominant kernel: the class launches (2 per sweep) or, for the chain schedule, one launch per k sweeps;
    # its own duration from HIP event pairs around each of its launches, over further sweeps after the region
    launches = steps * len(Bt.classes) if sch.kind == "class" else nrep
    bytes_per_launch = upd * steps / launches * 29.0        # 24 B spin read + write, 1 B mask, 4 B site K
    kms, kn = 0.0, 0
    for _ in range(5):
        m_, n_ = Bt.sweep_timed(S, k)
        kms, kn = kms + m_, kn + n_
    launch_s = kms / 1e3 / kn
    achieved = bytes_per_launch / launch_s / 1e9
    # CPU: the oracle's sequential chain on one replica of the median disk
    med = int(np.argsort([s.n_active for s in systems])[len(systems) // 2])
    os_ = O.OracleSystem(systems[med])
    beta = Bt.beta.cpu().numpy()[int(Bt.rep_off[med]):int(Bt.rep_off[med]) + reps]
    x = S0[med][[0]].astype(np.float32)
    t1 = time.perf_counter()
    sweeps = 0
    while time.perf_counter() - t1 < 12.0 and sweeps < 2000 and not a.no_cpu_baseline:
        x, _ = O.chain(os_, x, beta[:1], med, sweeps, 1, reps=[0])
        sweeps += 1
    dt = time.perf_counter() - t1
    out = {"metric": "site updates/s (%HBM roofline)", "value": value, "unit": "site updates/s", "n_gpus": 1,
           "steps": steps, "warmup": a.warmup, "ms_per_step": sec * 1e3 / steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (random initial spins from the seeded torch CPU generator)",
           "config": {"workload": WORKLOADS["nanodisks"], "systems": len(systems), "replicas_per_system": reps,
                      "active_sites": int(sum(s.n_active for s in systems)), "schedule": sch.label(),
                      "parallelism": "single GPU"},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles("nanodisks"),
                        "alg_bytes_per_launch": bytes_per_launch, "avg_launch_us": launch_s * 1e6,
                        "launches_timed": kn},
           "valu": valu_from_profiles("nanodisks"),
           "autotune": [[lbl, ms] for lbl, ms in Bt.tuning], "wall_s": wall}
    if not a.no_cpu_baseline and sweeps:
        ncores = min(16, reps, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1)
        xs = S0[med][:ncores].astype(np.float32)
        t2 = time.perf_counter()
        sw2 = 0
        while time.perf_counter() - t2 < 10.0 and sw2 < 2000:
            xs, _ = O.chain(os_, xs, beta, med, sw2, 1, reps=list(range(ncores)), threads=ncores)
            sw2 += 1
        dt2 = time.perf_counter() - t2
        out["cpu_baseline"] = {"value": systems[med].n_active * sweeps / dt, "unit": "site updates/s", "cores": 1,
                               "kind": "port", "cpu_model": _cpu_model(), "host_cores": _host_cores(),
                               "sample": f"1 replica of the median disk ({systems[med].n_active} sites) x {sweeps} "
                                         f"sweeps, sequential chain, {dt:.1f} s",
                               "all_cores": {"value": ncores * systems[med].n_active * sw2 / dt2, "cores": ncores,
                                             "sample": f"{ncores} replicas of the median disk (one per thread, OpenMP)"
                                                       f" x {sw2} sweeps, {dt2:.1f} s"}}
    print(json.dumps(out), flush=True)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="xy", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ess", action="store_true", help="default workload: skip the learned-vs-fixed comparison")
    ap.add_argument("--out-dir", default=None,
                    help="learned comparisons: also write the full comparison (per-temperature tau_int, z-scores) "
                         "as <workload>_comparison.json into this folder")
    ap.add_argument("--tile", default=None, help="replica-major kernels: class slots per tile, e.g. 8,8,4")
    ap.add_argument("--chunks", type=int, default=None, help="replica-major kernels: 64-replica chunks per wave")
    ap.add_argument("--resident-threads", type=int, default=None, help="LDS-resident kernel: threads per replica")
    ap.add_argument("--tile-dims", default=None, help="tiled sweep kernel: cells per tile along x and y, and optionally threads per workgroup, e.g. 64,64 or 128,32,512")
    ap.add_argument("--batch", type=int, default=None, help="learned-xy: replicas (default 512; 16 temperatures)")
    ap.add_argument("--layout", default="auto", choices=("auto", "reference", "replica"),
                    help="spin storage (replica: lanes over replicas, per-class launches)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy three times (the reference for the HBM counters)")
    ap.add_argument("--calibrate-cache-copy", action="store_true",
                    help="only run a plain 96 MiB device copy, whose source and destination stay in the Infinity Cache")
    ap.add_argument("--schedule", default="auto",
                    help="nanodisks: 'auto' (autotune) or kind/block/tile/order/grid, e.g. class/256/256/xcd/0")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "bench.py measures one GPU: the sweep path has no multi-GPU component "
                          "(replica blocks shard across GPUs without communication); run with --gpus 1"}))
        sys.exit(2)
    import torch

    from sck_amd.sampler import Sampler
    if a.calibrate_copy:
        x = torch.ones(2 ** 28, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        y.copy_(x)
        e0.record()
        for _ in range(3):
            y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 3
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": x.numel() * 4, "copies": 4,
                          "ms_per_copy": ms, "copy_GBps": 2 * x.numel() * 4 / (ms / 1e3) / 1e9}))
        return
    if a.calibrate_cache_copy:
        # the plain copy on 96 MiB: source and destination (192 MiB) stay in the 256 MB Infinity Cache, beyond
        # the 8 x 4 MB of L2, which is the level the replica-blocked and batched kernels stream from
        x = torch.ones(24 << 20, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for _ in range(5):
            y.copy_(x)
        e0.record()
        for _ in range(50):
            y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 50
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": x.numel() * 4, "copies": 55,
                          "ms_per_copy": ms, "cache_copy_GBps": 2 * x.numel() * 4 / (ms / 1e3) / 1e9}))
        return
    if a.workload in ("learned", "learned-xy"):
        learned_main(a)
        return
    if a.workload == "reference-table":
        reference_table_main(a)
        return
    if a.workload == "nanodisks":
        nanodisk_main(a)
        return
    s, kw, B, T, alg_bytes = _make(a.workload)
    if a.tile:
        kw["tile"] = tuple(int(v) for v in a.tile.split(","))
    if a.chunks:
        kw["chunks"] = a.chunks
    if a.resident_threads:
        kw["resident_threads"] = a.resident_threads
    if a.tile_dims:
        kw["tile_dims"] = tuple(int(v) for v in a.tile_dims.split(","))
    if a.layout != "auto":
        kw["layout"] = a.layout
        kw["tiled"] = False if a.layout == "replica" else None
    A = Sampler(s, B, "cuda:0", seed=0, **kw)
    A.set_temperature(T)
    if a.workload in ("xy", "heisenberg"):
        S = device_spins(A, seed=0)
    else:
        S = A.init_spins(seed=0)
    S0 = S[:16].cpu().numpy()                  # the replicas (words) the CPU baseline runs
    ncls = len(A.classes)
    upd_per_sweep = B * s.n_active
    resident = A.module.resident is not None and A.use_resident
    steps = max(1, a.steps)
    if a.warmup:
        A.sweep(S, a.warmup)
    # the timed region: one replay of a HIP graph of exactly `steps` sweeps (class launches and counter bumps)
    rp = A.graph(S, steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms = rp.time_ms(1)
    wall = time.perf_counter() - t0
    sec = ms / 1e3
    value = upd_per_sweep * steps / sec
    del rp
    # dominant kernel: the class launches (or the resident kernel). Its own duration comes from HIP event
    # pairs around each of its launches on the sampler's stream, over further sweeps right after the timed
    # region (the counter bump and launch gaps stay outside the pairs)
    # (a resident launch runs all `steps` sweeps of the timed graph, a class launch one class of one sweep)
    tsweeps = steps if resident else min(steps, 50)
    kms, kn = A.sweep_timed(S, tsweeps)
    launch_s = kms / 1e3 / kn
    bytes_per_launch = upd_per_sweep * tsweeps / kn * alg_bytes
    achieved = bytes_per_launch / launch_s / 1e9
    out = {
        "metric": "site updates/s (%HBM roofline)",
        "value": value,
        "unit": "site updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": ms / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bitwise": "u32", "ising": "int8", "heisenberg": "f32", "xy": "f32"}[a.workload],
        "data": "synthetic (random initial spins: seeded torch generator)",
        "config": {"workload": WORKLOADS[a.workload], "batch": B, "sites": s.n_active, "classes": ncls,
                   "kernel": "sck_resident" if resident else ("sck_tiled" if A.use_tiled else "sck_class_*"),
                   "layout": A.layout, "tile_dims": A.module.tile_dims if A.use_tiled else None,
                   "replica_blocks": A.module.rep_blocks, "timed": f"one replay of a HIP graph of {steps} sweeps",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles(a.workload),
                     "alg_bytes_per_launch": bytes_per_launch, "avg_launch_us": launch_s * 1e6,
                     "launches_timed": kn},
        "valu": valu_from_profiles(a.workload),
        "wall_s": wall,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.workload, s, A, S0)
    if a.workload == "xy" and not a.no_ess:
        del A, S
        torch.cuda.empty_cache()
        try:
            out["ess"] = ess_summary()
        except Exception as e:                 # the throughput line stands on its own
            out["ess"] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out, default=float), flush=True)


def ess_summary(log=None) -> dict:
    """Effective samples/s of the energy on the Heisenberg bcc 32^3 configuration (B = 8192, 32 temperatures in
    [0.5, 3]): the learned width over the best of five fixed widths and over the uniform redraw, with the
    per-temperature best fixed width (the ceiling of any width-only proposal) beside them."""
    import numpy as np

    from sck_amd.ess import compare_proposals
    s, _, B, T, _ = _make("heisenberg")
    temps = np.unique(T.numpy())
    r = compare_proposals(s, B, temps, warmup=200, equil=1000, therm=50, n_meas=1000,
                          log=log or (lambda m: print(m, file=sys.stderr, flush=True)))
    return {"metric": "effective samples/s, learned vs fixed proposal",
            "workload": WORKLOADS["heisenberg"],
            "protocol": "train on 200 fixed-width warm-up sweeps; 1000 learned-width sweeps to one common "
                        "equilibrium state; from it 50 + 1000 sweeps of each proposal with an observation after "
                        "every sweep; tau_int of E per temperature (Sokal window, c = 6)",
            "gain_E": r["gain_E"], "gain_E_vs_uniform": r["vs_uniform"]["gain_E"],
            "gain_E_vs_per_T_best_fixed": r["gain_E_vs_per_T_best"],
            "per_T_best_fixed_over_best_single_fixed": r["gain_E_per_T_best_over_fixed"],
            "gain_M": r["gain_M"], "fixed_width": r["fixed_width"],
            "observables_frac_within_2sigma": r["frac_within_2sigma"], "observables_max_z": r["max_z"],
            "ess_E_per_s": {k: r[k]["ess_E_per_s"] for k in ("uniform", "fixed", "learned")},
            "loop_updates_per_s": {k: r[k]["updates_per_s"] for k in ("uniform", "fixed", "learned")},
            "table_vs_fine_tau_E": r.get("table_vs_fine_tau_E"),
            "train_seconds": r["train_seconds"]}


if __name__ == "__main__":
    main()
