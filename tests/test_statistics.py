"""Long-chain agreement of the vector models with the oracle's sequential chain (the statistical bar for XY and
Heisenberg in fp32): E, |M|, Cv and chi of GPU chains against independent oracle chains -- a different Philox base
seed and a different initial state, so the two sides share no randomness -- at three temperatures each.

Systems: the Heisenberg bcc J1 = 1, J2 = -0.2, Du = 0.1 along z, h = 0.05 along z Hamiltonian on 4^3 cells, and
the XY model on the triangular lattice 12^2; the Heisenberg case also runs the cone proposal in the replica-major
layout (any proposal that keeps detailed balance samples the same distribution as the oracle's uniform redraw).
Per temperature: 64 GPU chains against 16 oracle chains, 500 sweeps discarded, then an observation after each of
2000 sweeps. Errors are delete-one-block jackknife errors over chains (SCK/simulate.py:521-526 defines Cv and chi;
the variances are pooled over the chains of a temperature). Bar: no difference is significant at family-wise
alpha = 0.05 after Holm's step-down correction over the 12 comparisons of a system (two-sided normal p-values of
|a - b| / sqrt(err_a^2 + err_b^2)).
"""

import math

import numpy as np
import pytest
import torch

import oracle as O
from sck_amd import crystals as cr
from sck_amd import system as sy

DEV = "cuda"
THERM, NMEAS = 500, 2000


def _heis():
    return sy.System("Heis bcc J1J2 D h", 3, cr.make("bcc", (1, 2)), (4, 4, 4), (True,) * 3, {1: 1.0, 2: -0.2},
                     h=(0.0, 0.0, 0.05), uniaxial=(0.1, (0.0, 0.0, 1.0)))


CASES = {
    "heisenberg_uniform": (_heis, (1.0, 1.9, 3.0), dict()),
    "heisenberg_cone_replica_major": (_heis, (1.0, 1.9, 3.0), dict(proposal="cone", width=1.0, layout="replica")),
    "xy_triangular_uniform": (lambda: sy.xy("triangular", 12), (0.8, 1.4, 2.2), dict()),
}


def _oracle_series(s, temps, chains, seed):
    """E (NMEAS, chains x temps) and |M| per site of independent oracle chains from random unit vectors."""
    os_ = O.OracleSystem(s)
    rng = np.random.default_rng(seed)
    C = chains * len(temps)
    x = rng.normal(size=(C, s.n) + s.shape)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    x = (x * s.site_mask()[None, None]).astype(np.float32)
    beta = (1.0 / np.repeat(np.asarray(temps, np.float32), chains)).astype(np.float32)
    x, _ = O.chain(os_, x, beta, seed, 0, THERM, threads=8)
    E = np.zeros((NMEAS, C))
    Ma = np.zeros((NMEAS, C))
    N = float(s.n_active)
    for t in range(NMEAS):
        x, _ = O.chain(os_, x, beta, seed, THERM + t, 1, threads=8)
        e, m = O.observe(os_, x)
        E[t] = e
        Ma[t] = np.linalg.norm(m.sum(axis=1), axis=1) / N
    return E, Ma


def _stats(E, Ma, temps, per, N):
    from sck_amd.ess import _stats
    groups = np.repeat(np.arange(len(temps)), per)
    return _stats(E, Ma, groups, np.asarray(temps), N)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_vector_model_observables_agree_with_oracle_chains(name):
    from sck_amd.observe import series
    from sck_amd.sampler import Sampler
    from sck_amd.validate import holm
    make, temps, kw = CASES[name]
    s = make()
    print(f"[gpu] statistics {name}: {len(temps)} temperatures, {THERM} + {NMEAS} sweeps", flush=True)
    gper, oper = 64, 16
    A = Sampler(s, gper * len(temps), DEV, seed=101, **kw)
    assert A.replica_major == (kw.get("layout") == "replica")
    A.set_temperature(torch.as_tensor(np.repeat(np.asarray(temps, np.float32), gper)))
    S = A.init_spins(seed=7)
    A.sweep(S, THERM)
    Eg, Mg, _ = series(A, S, NMEAS)
    Eo, Mo = _oracle_series(s, temps, oper, seed=2024)
    N = float(s.n_active)
    g, o = _stats(Eg, Mg, temps, gper, N), _stats(Eo, Mo, temps, oper, N)
    z, rows = [], []
    for a, b in zip(g, o):
        for k in ("E", "M", "Cv", "chi"):
            err = math.hypot(a[k + "_err"], b[k + "_err"])
            z.append(abs(a[k] - b[k]) / err)
            rows.append(f"T={a['T']:.2f} {k:3s} gpu {a[k]:+.5f} +- {a[k + '_err']:.5f}  oracle {b[k]:+.5f} +- "
                        f"{b[k + '_err']:.5f}  z {z[-1]:.2f}")
    print("\n".join(rows), flush=True)
    p = [math.erfc(v / math.sqrt(2.0)) for v in z]
    adj = holm(p)
    assert min(adj) > 0.05, "\n".join(rows)
    # the comparison has power: the energy per site is pinned to better than 1 % at every temperature
    assert all(r["E_err"] < 0.01 * abs(r["E"]) for r in o)
