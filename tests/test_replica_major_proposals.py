"""Cone and learned proposals in the replica-major layout (lanes over replicas, [b][L3][L2][L1][n][B] storage):
the class kernels behind the learned-vs-fixed comparisons at the Heisenberg bcc 32^3 and XY triangular 2048^2
configurations.

Each check forces layout="replica" and asserts the sampler took it:
  * every class's decisions against the oracle's sequential chain from the same state (B = 128), equal except
    where |beta dE + ln u| < 1e-5 (the fp32 decision tolerance);
  * the full Heisenberg bcc 32^3 batch (B = 8192, 32 replica blocks of 256) for three replicas per class;
  * one spin in a field, whose distribution is known in closed form, sampled in the replica-major layout (the
    single-site tests elsewhere run the LDS-resident kernel);
  * the sck_sample training records against the local field computed on the host from the same spins.
"""

import math

import numpy as np
import pytest
import torch

import oracle as O
from sck_amd import crystals as cr
from sck_amd import system as sy

DEV = "cuda"
pytestmark = pytest.mark.gpu


def _heis(box=(4, 4, 4)):
    return sy.System("Heis bcc J1J2 D h", 3, cr.make("bcc", (1, 2)), box, (True,) * 3, {1: 1.0, 2: -0.2},
                     h=(0.0, 0.0, 0.05), uniaxial=(0.1, (0.0, 0.0, 1.0)))


def _proposal_kwargs(kind, s, B, seed=7):
    from types import SimpleNamespace

    from sck_amd.proposal import LearnedWidth, init_params, mlp_config
    if kind == "cone":
        return dict(proposal="cone", width=np.linspace(0.2, 3.0, B).astype(np.float32)), None
    cfg = mlp_config(SimpleNamespace(sys=s))
    cfg["xs"] = (0.1, 1.0, 1.0)
    lw = LearnedWidth(init_params(9), cfg, n=(32, 8, 8), s=(2.0, 8.0, 8.0))
    return dict(proposal="learned", learned=lw), lw


def _oracle_prop(kind, A, lw):
    w = A.width.cpu().numpy()
    return O.cone(w) if kind == "cone" else O.learned(w, lw.n, lw.s, lw.cfg["axis"])


def _class_by_class(A, S, os_, prop, seed, reps=None, tol=2e-5):
    """Each class of sweep 0 from the current state, against the oracle; the compared replicas then continue
    from the oracle's state."""
    beta = A.beta.cpu().numpy()
    changed = 0
    for c in range(len(A.classes)):
        x0 = (S if reps is None else S[reps]).cpu().numpy()
        A.state[1] = 0
        A.sweep(S, 1, classes=(c, c + 1))
        ref, _, mg = O.chain(os_, np.ascontiguousarray(x0), beta, seed, 0, 1, reps=reps, prop=prop,
                             class_range=(c, c + 1), margins=True, threads=4)
        got = (S if reps is None else S[reps]).cpu().numpy()
        bad = np.abs(got - ref).max(axis=1) > tol
        near = np.abs(mg.reshape(bad.shape)) < 1e-5
        assert not np.any(bad & ~near), f"class {c}: {int((bad & ~near).sum())} decisions differ"
        changed += int((np.abs(got - x0).max(axis=1) > 0).sum())
        if reps is None:
            S.copy_(torch.as_tensor(ref, device=DEV))
        else:
            S[reps] = torch.as_tensor(ref, device=DEV)
    assert changed > 0


@pytest.mark.parametrize("model", ["heisenberg", "xy"])
@pytest.mark.parametrize("kind", ["cone", "learned"])
def test_replica_major_decisions_match_oracle(kind, model):
    from sck_amd.sampler import Sampler
    s = _heis() if model == "heisenberg" else sy.xy("triangular", 12)
    B = 128
    print(f"[gpu] replica-major {kind} decisions {s.name} B={B}", flush=True)
    kw, lw = _proposal_kwargs(kind, s, B)
    A = Sampler(s, B, DEV, seed=7, layout="replica", **kw)
    assert A.replica_major and A.module.rep_blocks == 1
    A.set_temperature(torch.linspace(0.4, 3.0, B))
    S = A.init_spins(seed=5)
    assert not S.is_contiguous()               # the permuted view of replica-major storage
    _class_by_class(A, S, O.OracleSystem(s), _oracle_prop(kind, A, lw), 7)


@pytest.mark.parametrize("kind", ["cone", "learned"])
def test_full_size_heisenberg_blocks_match_oracle(kind):
    """Heisenberg bcc 32^3, J1 = 1, J2 = -0.2, Du = 0.1, h = 0.05, B = 8192 (32 temperatures in [0.5, 3]), in 32
    cache-resident replica blocks of 256: replicas 0, 4097 and 8191 (first, middle and last block) against the
    oracle for every class."""
    import bench
    from sck_amd.sampler import Sampler
    s, _, B, T, _ = bench._make("heisenberg")
    print(f"[gpu] full-size heisenberg {kind}: B={B}", flush=True)
    kw, lw = _proposal_kwargs(kind, s, B, seed=3)
    if kind == "cone":
        kw["width"] = np.tile(np.array([0.3, 0.9, 1.7, 2.6], np.float32), B // 4)
    A = Sampler(s, B, DEV, seed=3, **kw)
    assert A.replica_major and A.module.rep_blocks == 32
    A.set_temperature(T)
    S = bench.device_spins(A, seed=1)
    _class_by_class(A, S, O.OracleSystem(s), _oracle_prop(kind, A, lw), 3, reps=[0, B // 2 + 1, B - 1])


@pytest.mark.parametrize("model", ["heisenberg", "xy"])
@pytest.mark.parametrize("kind", ["uniform", "cone", "learned"])
def test_replica_major_single_spin_distribution(kind, model):
    """One spin in a field: Heisenberg <s_z> = int x e^{b(hx + Dx^2)} dx / Z over [-1, 1]; XY <s_x> = I1(bh)/I0(bh).
    B = 4096 replicas in the replica-major layout sample it with every proposal (detailed balance)."""
    from scipy.special import iv
    from sck_amd.sampler import Sampler
    if model == "heisenberg":
        h, D, T = 0.7, 0.5, 0.6
        s = sy.System("single spin", 3, cr.make("sc"), (1, 1, 1), (False,) * 3, {1: 1.0}, h=(0.0, 0.0, h),
                      uniaxial=(D, (0.0, 0.0, 1.0)))
        x = np.linspace(-1, 1, 200001)
        wgt = np.exp((h * x + D * x * x) / T)
        exact, comp = (x * wgt).sum() / wgt.sum(), 2
    else:
        h, T = 0.8, 0.5
        s = sy.System("single XY spin", 2, cr.make("sc"), (1, 1, 1), (False,) * 3, {1: 1.0}, h=(h, 0.0))
        exact, comp = float(iv(1, h / T) / iv(0, h / T)), 0
    B = 4096
    kw = {}
    if kind == "cone":
        kw = dict(proposal="cone", width=0.6)
    elif kind == "learned":
        from types import SimpleNamespace

        from sck_amd.proposal import LearnedWidth, init_params, mlp_config
        cfg = mlp_config(SimpleNamespace(sys=s))
        cfg["xs"] = (0.5, 1.0, 1.0)
        kw = dict(proposal="learned", learned=LearnedWidth(init_params(5), cfg, n=(16, 4, 4), s=(4.0, 2.0, 4.0)))
    A = Sampler(s, B, DEV, seed=9, layout="replica", resident=False, **kw)
    assert A.replica_major and A.module.resident is None
    A.set_temperature(T)
    S = A.init_spins(seed=1)
    A.sweep(S, 200)
    acc = []
    for _ in range(40):
        A.sweep(S, 5)
        acc.append(S[:, comp].double().mean().item())
    est = float(np.mean(acc))
    err = float(np.std(acc)) / math.sqrt(len(acc)) * 3 + 2e-3
    assert abs(est - exact) < err, (kind, model, est, exact)


@pytest.mark.parametrize("model", ["heisenberg", "xy"])
def test_replica_major_sample_records(model):
    """sck_sample in the replica-major layout: each record holds the spin, the local field of its neighbours
    plus the field (computed on the host in float64 by rolls from the same spins), beta and the site's constants."""
    from sck_amd.proposal import collect_samples
    from sck_amd.sampler import Sampler
    s = _heis() if model == "heisenberg" else sy.xy("triangular", 12)
    B = 128
    A = Sampler(s, B, DEV, seed=2, proposal="cone", width=1.0, layout="replica")
    assert A.replica_major
    A.set_temperature(torch.linspace(0.5, 2.5, B))
    S = A.init_spins(seed=4)
    A.sweep(S, 3)
    recs = collect_samples(A, S, 4000, seed=5).cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(5)
    active = np.flatnonzero(s.site_mask().reshape(-1))
    reps = rng.integers(0, B, size=4000)
    sites = active[rng.integers(0, active.size, size=4000)]
    X = S.contiguous().double()
    v = A.local_field(X).reshape(B, s.n, -1).cpu().numpy()
    spins = X.reshape(B, s.n, -1).cpu().numpy()
    h = np.zeros(s.n) if s.h is None else np.asarray(s.h, np.float64)
    for k in range(s.n):
        assert np.array_equal(recs[:, k], spins[reps, k, sites])
        assert np.allclose(recs[:, 3 + k], v[reps, k, sites] + h[k], atol=2e-5)
    assert np.allclose(recs[:, 6], A.beta.cpu().numpy()[reps], rtol=0)
