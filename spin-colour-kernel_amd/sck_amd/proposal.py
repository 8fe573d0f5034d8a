"""Cone proposals for XY/Heisenberg: a fixed half-angle, and a learned one (no counterpart in the reference,
whose proposal is a uniform redraw, SCK/generator.py:274-289).

A cone move draws the trial spin uniformly from the cap of half-angle w around the current spin (XY: the arc
phi +- w). With w fixed, or a function of quantities the site's own move cannot change, the proposal density is
symmetric and plain Metropolis keeps detailed balance exactly. The learned width is such a function: a
3 -> 32 -> 32 -> 1 tanh network of

    f0 = |v| beta,  f1 = (|Du| + |Ku_i| + |K1_i|) beta,  f2 = |v . e| / |v|

(v the local field of the neighbours, e the anisotropy axis). It is trained on the GPU by libsck's own forward
and backward kernels (csrc/mlp.hip) on records of the sampler's warm-up sweeps, maximising the mean log of the
expected squared jump distance per update (so cold and hot sites weigh alike), then frozen into a table over
quantised features that the sweep kernel reads.
"""

from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L

__all__ = ["LearnedWidth", "train_width", "collect_samples", "NPARAM", "init_params", "mlp_config",
           "grad_step", "best_fixed_width"]

H = 32
NPARAM = H * 3 + H + H * H + H + H + 1
SAMPLE_FLOATS = 12


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases (the output bias starts at w ~ 1 rad)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((3, H), (H, H), (H, 1)):
        lim = math.sqrt(6.0 / (fan_in + fan_out))
        parts.append(rng.uniform(-lim, lim, size=fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    P = np.concatenate(parts).astype(np.float32)
    assert P.size == NPARAM
    return P


def mlp_config(sampler, ndraw: int = 4, seed: int = 7, wmin: float = 0.05, xs=(0.125, 1.0, 1.0)) -> dict:
    s = sampler.sys
    du, axis = (s.uniaxial[0], s.uniaxial[1]) if s.uniaxial is not None else (0.0, (0.0, 0.0, 1.0)[:s.n])
    ax = list(axis) + [0.0] * (3 - len(axis))
    return dict(n=s.n, ndraw=int(ndraw), seed=int(seed), wmin=float(wmin), wmax=float(math.pi),
                k2=float(s.cubic_k2), du=float(du), axis=tuple(float(v) for v in ax), xs=tuple(float(v) for v in xs))


def _cfg_struct(cfg: dict) -> L.MlpConfig:
    c = L.MlpConfig()
    c.n, c.ndraw, c.seed = cfg["n"], cfg["ndraw"], cfg["seed"]
    c.wmin, c.wmax, c.k2, c.du = cfg["wmin"], cfg["wmax"], cfg["k2"], cfg["du"]
    for k in range(3):
        c.axis[k] = cfg["axis"][k]
        c.xs[k] = cfg["xs"][k]
    return c


def collect_samples(sampler, S: torch.Tensor, count: int, seed: int = 0) -> torch.Tensor:
    """Training records (count, 12) of random (replica, active site) pairs of the current spins, written by the
    module's sck_sample kernel."""
    rng = np.random.default_rng(seed)
    active = np.flatnonzero(sampler.sys.site_mask().reshape(-1))
    reps = rng.integers(0, sampler.B, size=count)
    sites = active[rng.integers(0, active.size, size=count)]
    rs = torch.as_tensor(np.stack([reps, sites], 1).astype(np.int32).reshape(-1), device=sampler.device)
    out = torch.zeros((count, SAMPLE_FLOATS), dtype=torch.float32, device=sampler.device)
    T = sampler._native(S)
    a = sampler._args(T)
    with torch.cuda.device(sampler.device):
        L.check(L.lib().sck_sample_sites(sampler._handle, C.byref(a), C.c_void_p(rs.data_ptr()), int(count),
                                         C.c_void_p(out.data_ptr()), C.c_void_p(torch.cuda.current_stream().cuda_stream)),
                "sample sites")
    return out


def grad_step(params: torch.Tensor, samples: torch.Tensor, cfg: dict, lr: float = 0.0, step: int = 1,
              m: torch.Tensor | None = None, v: torch.Tensor | None = None, workspace: torch.Tensor | None = None):
    """One GPU training step (lr = 0: gradient only). Returns grad (NPARAM + 1,): dL/dP with
    L = -mean_i log(1e-4 + J_i), and the summed objective log(1e-4 + J_i) last."""
    dev = params.device
    ns = int(samples.shape[0])
    need = int(L.lib().sck_mlp_workspace_bytes(ns, cfg["ndraw"]))
    if workspace is None or workspace.numel() * 4 < need:
        workspace = torch.empty((need + 3) // 4, dtype=torch.float32, device=dev)
    grad = torch.zeros(NPARAM + 1, dtype=torch.float32, device=dev)
    ac = L.AdamConfig()
    ac.lr, ac.beta1, ac.beta2, ac.eps, ac.step = float(lr), 0.9, 0.999, 1e-8, int(step)
    c = _cfg_struct(cfg)
    with torch.cuda.device(dev):
        L.check(L.lib().sck_mlp_train_step(C.c_void_p(params.data_ptr()),
                                           C.c_void_p(m.data_ptr() if m is not None else 0),
                                           C.c_void_p(v.data_ptr() if v is not None else 0),
                                           C.c_void_p(samples.data_ptr()), ns, C.byref(c), C.byref(ac),
                                           C.c_void_p(workspace.data_ptr()), C.c_void_p(grad.data_ptr()),
                                           C.c_void_p(torch.cuda.current_stream().cuda_stream)), "training step")
    return grad


@dataclass
class LearnedWidth:
    """A trained width network and its table over quantised features (the sweep kernel's lookup)."""

    params: np.ndarray
    cfg: dict
    n: tuple = (64, 16, 16)
    s: tuple = (4.0, 8.0, 16.0)
    history: list = field(default_factory=list)
    _table: np.ndarray | None = None

    def meta(self) -> dict:
        return dict(n=tuple(int(x) for x in self.n), s=tuple(float(x) for x in self.s), axis=tuple(self.cfg["axis"]))

    def table(self, beta=None) -> np.ndarray:
        """Width at every bin centre, flattened [i0][i1][i2] (independent of beta: beta enters the features)."""
        if self._table is None:
            dev = torch.device("cuda", torch.cuda.current_device())
            P = torch.as_tensor(self.params, device=dev)
            out = torch.zeros(int(np.prod(self.n)), dtype=torch.float32, device=dev)
            g = L.WidthGrid()
            for d in range(3):
                g.n[d], g.s[d] = int(self.n[d]), float(self.s[d])
            c = _cfg_struct(self.cfg)
            with torch.cuda.device(dev):
                L.check(L.lib().sck_mlp_table(C.c_void_p(P.data_ptr()), C.byref(c), C.byref(g),
                                              C.c_void_p(out.data_ptr()), C.c_void_p(torch.cuda.current_stream().cuda_stream)),
                        "width table")
            self._table = out.cpu().numpy()
        return self._table


def train_width(sampler, S: torch.Tensor, warmup: int = 200, snapshots: int = 10, per_snapshot: int = 8192,
                steps: int = 400, lr: float = 3e-3, batch: int = 16384, ndraw: int = 4, seed: int = 11,
                grid=(64, 16, 16), log=None) -> LearnedWidth:
    """Run `warmup` sweeps of `sampler` (a fixed-width cone sampler) on S, taking `per_snapshot` records at
    `snapshots` evenly spaced points, then train the width network on them with Adam on the GPU."""
    dev = sampler.device
    every = max(1, warmup // snapshots)
    recs = []
    done = 0
    while done < warmup:
        k = min(every, warmup - done)
        sampler.sweep(S, k)
        done += k
        recs.append(collect_samples(sampler, S, per_snapshot, seed=seed + done))
    samples = torch.cat(recs)
    cfg = mlp_config(sampler, ndraw=ndraw, seed=seed)
    # feature ranges of the data fix the table grid: f0 up to its 99.9th percentile, f1 and f2 as observed
    hv = samples[:, 3:3 + sampler.sys.n].norm(dim=1)
    f0 = (hv * samples[:, 6]).cpu().numpy()
    f1 = ((abs(cfg["du"]) + samples[:, 8].abs() + samples[:, 7].abs()) * samples[:, 6]).cpu().numpy()
    f0max = float(np.quantile(f0, 0.999)) * 1.25 + 1e-3
    f1max = float(f1.max()) * 1.25 + 1e-3
    cfg["xs"] = (1.0 / f0max, 1.0 / f1max, 1.0)
    n = tuple(int(x) for x in grid)
    s = (n[0] / f0max, n[1] / f1max, float(n[2]))
    P = torch.as_tensor(init_params(seed), device=dev)
    m = torch.zeros_like(P)
    v = torch.zeros_like(P)
    rng = np.random.default_rng(seed)
    hist = []
    ws = None
    t0 = time.perf_counter()
    for step in range(1, steps + 1):
        idx = torch.as_tensor(rng.choice(samples.shape[0], size=min(batch, samples.shape[0]), replace=False),
                              device=dev)
        cfg["seed"] = seed + step
        g = grad_step(P, samples[idx].contiguous(), cfg, lr=lr, step=step, m=m, v=v, workspace=ws)
        if step == 1 or step % 50 == 0 or step == steps:
            J = float(g[-1]) / min(batch, samples.shape[0])
            hist.append((step, J))
            if log:
                log(f"  width training step {step:4d}: mean log expected squared jump {J:.5f}")
    torch.cuda.synchronize()
    cfg["seed"] = seed
    lw = LearnedWidth(P.cpu().numpy(), cfg, n=n, s=s, history=hist)
    lw.train_seconds = time.perf_counter() - t0
    lw.samples = samples
    return lw


def best_fixed_width(sampler, samples: torch.Tensor, widths=(0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 3.14159), ndraw=4):
    """The single cone half-angle with the largest training objective (mean log expected squared jump) over
    the training records: the fixed-width baseline the learned width is judged against (the same objective
    restricted to one constant)."""
    best, bestJ = None, -math.inf
    cfg = mlp_config(sampler, ndraw=ndraw)
    dev = samples.device
    res = []
    for w in widths:
        # a network whose output is the constant logit of w: all weights zero, bias at the logit
        P = np.zeros(NPARAM, np.float32)
        q = (w - cfg["wmin"]) / (cfg["wmax"] - cfg["wmin"])
        q = min(max(q, 1e-6), 1 - 1e-6)
        P[-1] = math.log(q / (1 - q))
        g = grad_step(torch.as_tensor(P, device=dev), samples, cfg)
        J = float(g[-1]) / samples.shape[0]
        res.append((w, J))
        if J > bestJ:
            best, bestJ = w, J
    return best, res
