"""Distribution-level validation against exact enumeration (the reference's experiment E2).

CPU: the enumeration against the known density of states of the 4 x 4 periodic Ising model and against the
closed form of an open two-site chain in a field; bin merging and Holm's correction.
GPU: sampled energies of the coloured sweep pass a chi-square test against the exact distribution (int8,
packed and bitwise storage, Barker and Metropolis), and the synchronous all-sites control is rejected.
"""

import math

import numpy as np
import pytest
import torch

from sck_amd import crystals as cr
from sck_amd import system as sy
from sck_amd.validate import energy_test, exact_energy_distribution, holm, merge_bins

DEV = "cuda"


def test_square_4x4_density_of_states():
    ex = exact_energy_distribution(sy.ising("square", 4), 2.269)
    got = dict(zip(ex["levels"].astype(int).tolist(), ex["degeneracy"].tolist()))
    # 2^16 states: g(-32) = 2, g(-24) = 32, g(-20) = 64, g(-16) = 424, g(-12) = 1728, g(-8) = 6688,
    # g(-4) = 13568, g(0) = 20524, symmetric under E -> -E
    half = {-32: 2, -24: 32, -20: 64, -16: 424, -12: 1728, -8: 6688, -4: 13568}
    want = dict(half)
    want.update({-k: v for k, v in half.items()})
    want[0] = 20524
    assert got == want and sum(got.values()) == 2 ** 16


def test_two_spins_in_a_field_closed_form():
    s = sy.System("pair", 1, cr.make("sc"), (2, 1, 1), (False,) * 3, {1: 1.5}, h=(0.4,))
    T = 1.3
    ex = exact_energy_distribution(s, T)
    # states: ++ : -J - 2h, -- : -J + 2h, +- and -+ : +J
    E = {round(-1.5 - 0.8, 6): 1, round(-1.5 + 0.8, 6): 1, 1.5: 2}
    Z = sum(g * math.exp(-e / T) for e, g in E.items())
    for lv, p in zip(ex["levels"], ex["p"]):
        assert abs(p - E[round(float(lv), 6)] * math.exp(-lv / T) / Z) < 1e-12


def test_bin_merging_and_holm():
    o, e = merge_bins(np.array([1, 2, 10, 3, 1.0]), np.array([3, 3, 9, 2, 4.0]))
    assert np.array_equal(e, [6.0, 9.0, 6.0]) and np.array_equal(o, [3.0, 10.0, 4.0])
    o, e = merge_bins(np.array([1, 2, 10, 3, 1.0]), np.array([3, 3, 9, 2, 1.0]))      # short tail joins the last bin
    assert np.array_equal(e, [6.0, 12.0]) and np.array_equal(o, [3.0, 14.0])
    assert np.allclose(holm([0.01, 0.04, 0.03]), [0.03, 0.06, 0.06])


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["int8-barker-sq4", "packed-barker-sq4", "bitwise-barker-sq4", "int8-metropolis-sq4",
                                  "int8-barker-tri3"])
def test_coloured_sweep_samples_the_exact_distribution(case):
    """Metropolis on the 3 x 3 triangular lattice is left out on purpose: with the deterministic class scan every
    dE = 0 flip is certain and the chain is not ergodic there (the sequential oracle, the reference's chain,
    shows the same bias); the heat-bath (Barker) rule samples it."""
    from sck_amd.sampler import Sampler
    store, acc, lat = case.split("-")
    if lat == "tri3":
        s, T = sy.ising("triangular", 3), 4.0 / math.log(3.0)
    else:
        s, T = sy.ising("square", 4), 2.0 / math.log(1 + math.sqrt(2))
    s.acceptance = acc
    kw = {"int8": dict(packed=False), "packed": dict(packed=True), "bitwise": dict(bitwise=True, independent=True)}[
        store]
    print(f"[gpu] exact distribution {case}", flush=True)
    r = energy_test(lambda seed: Sampler(s, 1024, DEV, seed=seed, **kw), s, T, seeds=(1, 2, 3), samples=30000,
                    batch=1024)
    assert all(row["off_level"] == 0 for row in r["rows"])
    assert r["rejected"] == 0, r["rows"]


@pytest.mark.gpu
def test_synchronous_control_is_rejected():
    """The all-sites update is not a valid sampler of the Boltzmann distribution: the test must see it."""
    from sck_amd.sampler import Sampler
    s, T = sy.ising("square", 4), 2.0 / math.log(1 + math.sqrt(2))
    s.acceptance = "barker"
    r = energy_test(lambda seed: Sampler(s, 1024, DEV, seed=seed, packed=False), s, T, seeds=(1, 2, 3),
                    samples=30000, batch=1024, rule="all sites")
    assert r["rejected"] == 3, r["rows"]


@pytest.mark.gpu
def test_rate_of_graph_replays():
    from sck_amd.perf import boot_ci, rate
    r, info = rate(sy.ising("square", 32), 256, DEV, T=2.5, reps=5)
    assert len(r) == 5 and min(r) > 0 and info["sweeps_per_replay"] == math.ceil(2e7 / (256 * 1024))
    m, lo, hi = boot_ci(r)
    assert lo <= m <= hi


# ------------------------------------------------------------------------------ application observables
def test_topological_charge_of_uniform_and_skyrmion_states():
    from sck_amd.observables import topological_charge
    L = 41
    s = sy.System("heis square", 3, cr.make("square"), (L, L), (False, False), {1: 1.0})
    yy, xx = np.meshgrid(np.arange(L) - L // 2, np.arange(L) - L // 2, indexing="ij")
    lam = 5.0
    r2 = xx ** 2 + yy ** 2
    # Belavin-Polyakov skyrmion: m = (2 lam x, 2 lam y, lam^2 - r^2) / (lam^2 + r^2), m_z = +1 at the core
    sk = np.stack([2 * lam * xx, 2 * lam * yy, lam ** 2 - r2]) / (lam ** 2 + r2)
    up = np.zeros_like(sk)
    up[2] = 1.0
    S = torch.as_tensor(np.stack([up, sk, -sk])[:, :, None, None], dtype=torch.float32)   # (3, 3, 1, 1, L, L)
    Q = topological_charge(S, s).numpy()
    assert abs(Q[0]) < 1e-12
    # the lattice charge of a finite box misses the tail beyond it: |Q| = 1 - lam^2 / (lam^2 + R^2) roughly
    assert 0.9 < abs(Q[1]) < 1.0 and abs(Q[1] + Q[2]) < 1e-9


def test_structure_factor_of_the_120_degree_state():
    from sck_amd.observables import site_positions, structure_factor
    L = 12
    s = sy.xy("triangular", L)
    K = np.array([4 * np.pi / 3, 0.0])
    pos = site_positions(s)
    th = pos @ np.pad(K, (0, 1))                                   # the 120 degree state: angle K . r
    S = torch.as_tensor(np.stack([np.cos(th), np.sin(th)])[None], dtype=torch.float64)
    sk = structure_factor(S, s, K).numpy()
    assert abs(sk[0] - L * L / 2) < 1e-9
    ferro = torch.zeros_like(S)
    ferro[:, 0] = 1.0
    assert abs(structure_factor(ferro, s, [0.0, 0.0]).numpy()[0] - L * L) < 1e-9


# ------------------------------------------------------------------------------ multi-histogram reweighting
def _exact_samples(s, temps, n, seed):
    """Independent canonical energy samples of an Ising system at each temperature, from its exact density
    of states (no Markov chain), with the exact levels and degeneracies."""
    rng = np.random.default_rng(seed)
    ref = exact_energy_distribution(s, temps[0])
    lv, g = ref["levels"], ref["degeneracy"].astype(float)
    out = []
    for T in temps:
        w = g * np.exp(-(lv - lv.min()) / T)
        out.append(rng.choice(lv, size=n, p=w / w.sum()))
    return np.array(out), lv, g


def test_multi_histogram_reproduces_exact_canonical_averages():
    """Samples at T = 2.0, 2.3, 2.6, 3.0 of the 4x4 Ising model reweighted to temperatures between and
    beyond them agree with <E> and <E^2> from exact enumeration; the free energies agree with exact ln Z."""
    from sck_amd.reweight import multi_histogram
    s = sy.ising("square", 4)
    temps = np.array([2.0, 2.3, 2.6, 3.0])
    E, lv, g = _exact_samples(s, temps, 40000, seed=3)
    grid = np.array([2.0, 2.15, 2.45, 2.8, 3.0])
    (mE, mE2), lnZ = multi_histogram(E, [E, E * E], temps, grid, return_lnZ=True)
    for T, a, a2 in zip(grid, mE, mE2):
        w = g * np.exp(-(lv - lv.min()) / T)
        p = w / w.sum()
        ex, ex2 = float(p @ lv), float(p @ lv ** 2)
        se = math.sqrt(ex2 - ex * ex) / math.sqrt(40000)
        assert abs(a - ex) < 5 * se, (T, a, ex, se)
        assert abs(a2 - ex2) < 5 * se * 2 * abs(ex) + 1e-9, (T, a2, ex2)
    e0 = E.mean()
    lz = np.array([np.log(np.sum(g * np.exp(-(lv - e0) / T))) for T in temps])
    assert np.allclose(lnZ, lz - lz[0], atol=0.02), (lnZ, lz - lz[0])


def test_multi_histogram_identities():
    """One temperature evaluated at itself gives the sample mean; a weight of 2 equals a duplicated sample;
    a zero weight drops a sample."""
    from sck_amd.reweight import multi_histogram
    rng = np.random.default_rng(1)
    E = rng.normal(-20.0, 3.0, (1, 500))
    O = rng.normal(0.0, 1.0, (1, 500))
    (m,) = multi_histogram(E, [O], [2.5], [2.5])
    assert abs(m[0] - O.mean()) < 1e-12
    E2 = rng.normal(-20.0, 3.0, (2, 300))
    O2 = E2 ** 2
    w = np.ones_like(E2)
    w[0, 5] = 2.0
    w[1, 7] = 0.0
    a = multi_histogram(E2, [O2], [2.0, 2.4], [2.2], weights=w)[0]
    Ed = [np.concatenate([E2[0], E2[0, 5:6]]), np.delete(E2[1], 7)]
    n = max(len(r) for r in Ed)
    pad = lambda r: np.concatenate([r, np.zeros(n - len(r))])                   # noqa: E731
    wd = np.array([pad(np.ones(len(r))) for r in Ed])
    Epad = np.array([pad(r) for r in Ed])
    b = multi_histogram(Epad, [Epad ** 2], [2.0, 2.4], [2.2], weights=wd)[0]
    assert abs(a[0] - b[0]) < 1e-9 * abs(b[0])
    with pytest.raises(ValueError):
        multi_histogram(E2, [O2[:, :3]], [2.0, 2.4], [2.2])
