"""Learned proposal width and the statistics helpers.

CPU: the oracle's analytic gradient of the training objective against central finite differences of the same
objective (pins the backpropagation and the cone derivative), tau_int known answers.
GPU: the training kernels' gradient against the oracle's, training raising the objective, the cone and learned
proposals against the oracle's decisions, and detailed balance of both on a single spin whose distribution is
known in closed form.
"""

import math

import numpy as np
import pytest
import torch

import oracle as O
from oracle import mlp as OM
from sck_amd import crystals as cr
from sck_amd import system as sy
from sck_amd.observe import integrated_tau, jackknife


def _fake_samples(n=48, seed=0, dims=3):
    rng = np.random.default_rng(seed)
    s = rng.normal(size=(n, 3))
    if dims == 2:
        s[:, 2] = 0
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    v = rng.normal(size=(n, 3)) * 2.0
    if dims == 2:
        v[:, 2] = 0
    beta = rng.uniform(0.3, 2.0, size=n)
    k1 = rng.uniform(0, 0.3, size=n) if dims == 3 else np.zeros(n)
    ku = rng.uniform(0, 0.3, size=n) if dims == 3 else np.zeros(n)
    ax = rng.normal(size=(n, 3))
    if dims == 2:
        ax[:, 2] = 0
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    return np.concatenate([s, v, beta[:, None], k1[:, None], ku[:, None], ax], 1).astype(np.float32)


def _cfg(n=3):
    return dict(n=n, ndraw=3, seed=5, wmin=0.05, wmax=math.pi, k2=0.05 if n == 3 else 0.0, du=0.1,
                axis=(0.0, 0.0, 1.0) if n == 3 else (1.0, 0.0, 0.0), xs=(0.2, 1.0, 1.0))


@pytest.mark.parametrize("n", [2, 3])
def test_oracle_gradient_matches_finite_differences(n):
    from sck_amd.proposal import init_params
    S = _fake_samples(40, seed=n, dims=n)
    cfg = _cfg(n)
    P = init_params(3).astype(np.float64)
    J, g = OM.objective_and_grad(P, S, cfg)
    assert np.isfinite(J)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(OM.NPARAM, 12, replace=False), [OM.NPARAM - 1, OM.OFF_B1, OM.OFF_W3]])
    eps = 1e-5
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += eps
        Pm[i] -= eps
        fd = (OM.loss(Pp, S, cfg) - OM.loss(Pm, S, cfg)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_integrated_tau_known_answers():
    rng = np.random.default_rng(0)
    assert integrated_tau(np.ones((100, 4))) == 1.0
    phi = 0.8
    x = np.zeros((20000, 32))
    for t in range(1, x.shape[0]):
        x[t] = phi * x[t - 1] + rng.normal(size=32)
    tau = integrated_tau(x)
    assert abs(tau - (1 + phi) / (1 - phi)) / ((1 + phi) / (1 - phi)) < 0.08
    w = integrated_tau(rng.normal(size=(5000, 16)))
    assert 0.8 < w < 1.2


def test_pooled_variance_is_free_of_the_autocorrelation_bias():
    """AR(1) chains x_t = phi x_{t-1} + eps of known variance 1 / (1 - phi^2): the per-chain estimator (the
    reference's) is low by about 2 tau / n, the pooled one is not."""
    from sck_amd.observe import cv_chi
    rng = np.random.default_rng(0)
    phi, n, c = 0.9, 400, 256
    x = np.zeros((n, c))
    x[0] = rng.normal(size=c) / math.sqrt(1 - phi * phi)
    for t in range(1, n):
        x[t] = phi * x[t - 1] + rng.normal(size=c)
    true = 1 / (1 - phi * phi)
    (per, per_e), _ = cv_chi(x, x, 1.0, 1.0)
    (pool, pool_e), (chi, _) = cv_chi(x, x, 1.0, 1.0, pooled=True)
    assert abs(pool - true) < 3 * pool_e and chi == pool
    assert per < true - 2 * per_e


def test_jackknife_error_of_the_mean():
    rng = np.random.default_rng(1)
    x = rng.normal(size=4000)
    m, e = jackknife(np.mean, x, blocks=40)
    assert abs(e - 1 / math.sqrt(4000)) < 0.3 / math.sqrt(4000)


# ------------------------------------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("n", [2, 3])
def test_gpu_gradient_matches_oracle(n):
    from sck_amd.proposal import NPARAM, grad_step, init_params
    S = _fake_samples(300, seed=10 + n, dims=n)
    cfg = _cfg(n)
    P = init_params(4)
    g = grad_step(torch.as_tensor(P, device="cuda"), torch.as_tensor(S, device="cuda"), cfg).cpu().numpy()
    J, go = OM.objective_and_grad(P.astype(np.float64), S, cfg)
    assert abs(g[-1] / 300 - J) < 1e-4 * max(1.0, abs(J))
    rel = np.linalg.norm(g[:NPARAM] - go) / np.linalg.norm(go)
    assert rel < 2e-3, rel


def _heis_system():
    return sy.System("Heis bcc J1J2 D h", 3, cr.make("bcc", (1, 2)), (4, 4, 4), (True,) * 3, {1: 1.0, 2: -0.2},
                     h=(0.0, 0.0, 0.05), uniaxial=(0.1, (0.0, 0.0, 1.0)))


@pytest.mark.gpu
def test_training_raises_the_objective_and_beats_the_best_fixed_width():
    from sck_amd.proposal import best_fixed_width, train_width
    from sck_amd.sampler import Sampler
    s = _heis_system()
    B = 256
    A = Sampler(s, B, "cuda", seed=3, proposal="cone", width=1.0)
    A.set_temperature(torch.as_tensor(np.repeat(np.linspace(0.5, 3.0, 8).astype(np.float32), B // 8)))
    S = A.init_spins(seed=2)
    lw = train_width(A, S, warmup=60, snapshots=6, per_snapshot=2048, steps=150, batch=8192)
    J0, J1 = lw.history[0][1], lw.history[-1][1]          # mean log expected squared jump
    assert J1 > J0 + 0.05
    wbest, res = best_fixed_width(A, lw.samples)
    assert J1 > max(J for _, J in res)
    tab = lw.table()
    assert tab.shape == (int(np.prod(lw.n)),) and np.all(tab > 0.04) and np.all(tab <= math.pi + 1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["heisenberg", "xy"])
@pytest.mark.parametrize("kind", ["cone", "learned"])
def test_cone_and_learned_decisions_match_oracle(kind, model):
    from sck_amd.proposal import LearnedWidth, init_params, mlp_config
    from sck_amd.sampler import Sampler
    s = _heis_system() if model == "heisenberg" else sy.xy("triangular", 12)
    print(f"[gpu] {kind} decisions {s.name}", flush=True)
    B = 16
    kw = {}
    if kind == "cone":
        kw = dict(proposal="cone", width=np.linspace(0.2, 3.0, B).astype(np.float32))
    else:
        probe = Sampler(s, B, "cuda", seed=7, proposal="cone", width=1.0)
        cfg = mlp_config(probe)
        cfg["xs"] = (0.1, 1.0, 1.0)
        lw = LearnedWidth(init_params(9), cfg, n=(32, 8, 8), s=(2.0, 8.0, 8.0))
        kw = dict(proposal="learned", learned=lw)
    A = Sampler(s, B, "cuda", seed=7, **kw)
    A.set_temperature(torch.linspace(0.4, 3.0, B))
    S = A.init_spins(seed=5)
    os_ = O.OracleSystem(s)
    beta = A.beta.cpu().numpy()
    prop = O.cone(A.width.cpu().numpy()) if kind == "cone" else \
        O.learned(A.width.cpu().numpy(), lw.n, lw.s, lw.cfg["axis"])
    for c in range(len(A.classes)):
        x0 = S.cpu().numpy()
        A.state[1] = 0
        A.sweep(S, 1, classes=(c, c + 1))
        ref, _, mg = O.chain(os_, x0, beta, 7, 0, 1, prop=prop, class_range=(c, c + 1), margins=True)
        got = S.cpu().numpy()
        bad = np.abs(got - ref).max(axis=1) > 2e-5
        near = np.abs(mg.reshape(bad.shape)) < 1e-5
        assert not np.any(bad & ~near), f"class {c}: {int((bad & ~near).sum())} decisions differ"
        S.copy_(torch.as_tensor(ref, device="cuda"))


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["uniform", "cone", "learned"])
def test_single_spin_distribution_is_exact(kind):
    """One Heisenberg spin in a field h along z with uniaxial D: <s_z> = int x e^{b(hx + Dx^2)} dx / Z over
    x in [-1, 1]; every proposal must sample it (detailed balance of the symmetric cone moves)."""
    from sck_amd.proposal import LearnedWidth, init_params, mlp_config
    from sck_amd.sampler import Sampler
    h, D, T = 0.7, 0.5, 0.6
    s = sy.System("single spin", 3, cr.make("sc"), (1, 1, 1), (False,) * 3, {1: 1.0}, h=(0.0, 0.0, h),
                  uniaxial=(D, (0.0, 0.0, 1.0)))
    B = 4096
    kw = {}
    if kind == "cone":
        kw = dict(proposal="cone", width=0.6)
    elif kind == "learned":
        probe = Sampler(s, 64, "cuda", seed=1, proposal="cone", width=1.0)
        cfg = mlp_config(probe)
        cfg["xs"] = (0.5, 1.0, 1.0)
        lw = LearnedWidth(init_params(5), cfg, n=(16, 4, 4), s=(4.0, 2.0, 4.0))
        kw = dict(proposal="learned", learned=lw)
    A = Sampler(s, B, "cuda", seed=9, **kw)
    A.set_temperature(T)
    S = A.init_spins(seed=1)
    A.sweep(S, 200)
    acc = []
    for _ in range(40):
        A.sweep(S, 5)
        acc.append(S[:, 2].double().mean().item())
    x = np.linspace(-1, 1, 200001)
    wgt = np.exp((h * x + D * x * x) / T)
    exact = (x * wgt).sum() / wgt.sum()
    est = float(np.mean(acc))
    err = float(np.std(acc)) / math.sqrt(len(acc)) * 3 + 2e-3
    assert abs(est - exact) < err, (kind, est, exact)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["uniform", "cone", "learned"])
def test_single_xy_spin_distribution_is_exact(kind):
    """One XY spin in a field h along x: <s_x> = I1(beta h) / I0(beta h) (modified Bessel functions); every
    proposal must sample it."""
    from scipy.special import iv
    from sck_amd.proposal import LearnedWidth, init_params, mlp_config
    from sck_amd.sampler import Sampler
    h, T = 0.8, 0.5
    s = sy.System("single XY spin", 2, cr.make("sc"), (1, 1, 1), (False,) * 3, {1: 1.0}, h=(h, 0.0))
    B = 4096
    kw = {}
    if kind == "cone":
        kw = dict(proposal="cone", width=0.9)
    elif kind == "learned":
        probe = Sampler(s, 64, "cuda", seed=1, proposal="cone", width=1.0)
        cfg = mlp_config(probe)
        cfg["xs"] = (0.5, 1.0, 1.0)
        lw = LearnedWidth(init_params(6), cfg, n=(16, 4, 4), s=(4.0, 2.0, 4.0))
        kw = dict(proposal="learned", learned=lw)
    A = Sampler(s, B, "cuda", seed=9, **kw)
    A.set_temperature(T)
    S = A.init_spins(seed=1)
    A.sweep(S, 200)
    acc = []
    for _ in range(40):
        A.sweep(S, 5)
        acc.append(S[:, 0].double().mean().item())
    exact = float(iv(1, h / T) / iv(0, h / T))
    est = float(np.mean(acc))
    err = float(np.std(acc)) / math.sqrt(len(acc)) * 3 + 2e-3
    assert abs(est - exact) < err, (kind, est, exact)


@pytest.mark.gpu
def test_learned_width_gains_at_the_cold_end():
    """Heisenberg bcc J1-J2 + Du + h on 8^3 cells, B = 256 at T = 0.5 and 0.6: the trained width decorrelates the
    energy faster than the uniform redraw (the reference's proposal) at the coldest temperature, where small moves
    are accepted and large ones are not: tau_int(E) of the learned width at most 0.75 of the uniform redraw's
    (measured 0.4 at the Heisenberg bcc 32^3 configuration, where the ESS/s gain at T = 0.5 is 2.5x). The
    observables of both agree within the jackknife errors."""
    from sck_amd.ess import compare_proposals
    s = sy.System("Heis bcc J1J2 D h", 3, cr.make("bcc", (1, 2)), (8, 8, 8), (True,) * 3, {1: 1.0, 2: -0.2},
                  h=(0.0, 0.0, 0.05), uniaxial=(0.1, (0.0, 0.0, 1.0)))
    r = compare_proposals(s, 256, [0.5, 0.6], warmup=100, equil=300, therm=20, n_meas=400, train_steps=200,
                          widths=(0.5, 1.0), log=lambda m: print(m, flush=True))
    t_u, t_l = r["uniform"]["tau_E"][0], r["learned"]["tau_E"][0]
    print(f"tau_int(E) at T = 0.5: uniform {t_u:.2f}, learned {t_l:.2f}", flush=True)
    assert t_l < 0.75 * t_u
    assert r["max_z"] < 4.0
