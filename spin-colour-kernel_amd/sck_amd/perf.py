"""Rate measurement of the generated kernels (the counterpart of SCK/exp_perf.py:22-57).

`rate` captures k = ceil(target_updates / (B N)) sweeps in one HIP graph, warms it up with two replays and
times `reps` single replays with HIP events on the replay's stream; it returns the per-replay rates in spin
updates per ns, as the reference's `rate` does. `boot_ci` is the bootstrap interval used to report them.
"""

from __future__ import annotations

import math

import numpy as np

from .sampler import Sampler
from .system import System

__all__ = ["rate", "boot_ci"]


def boot_ci(x, stat=np.mean, n: int = 2000, seed: int = 0, level: float = 0.95):
    """(stat(x), lower, upper) with a percentile bootstrap interval."""
    rng = np.random.default_rng(seed)
    x = np.asarray(x)
    bs = np.array([stat(x[rng.integers(0, len(x), len(x))]) for _ in range(n)])
    lo, hi = np.quantile(bs, [(1 - level) / 2, 1 - (1 - level) / 2])
    return float(stat(x)), float(lo), float(hi)


def rate(s: System, B: int, device="cuda", T: float | None = None, target_updates: float = 2e7, reps: int = 10,
         seed: int = 1, **sampler_kw):
    """Spin updates per ns over `reps` timed graph replays; returns (per-replay rates, info)."""
    A = Sampler(s, B, device, seed, **sampler_kw)
    A.set_temperature(T or s.T_ref or 1.0)
    S = A.init_spins(seed=seed)
    per_sweep = B * s.n_active
    k = max(1, int(math.ceil(target_updates / per_sweep)))
    replay = A.graph(S, k)
    for _ in range(2):
        replay()
    out = [per_sweep * k / (replay.time_ms(1) * 1e6) for _ in range(reps)]
    return out, {"sweeps_per_replay": k, "updates_per_sweep": per_sweep, "layout": A.layout}
