"""The nanodot cooling ladder (sck_amd.anneal; reference SCK/anneal.py): parameters, the accumulator reduction,
the command line, and on the GPU the graph-driven ladder against the same chain driven eagerly."""

import json

import numpy as np
import pytest
import torch

from sck_amd import anneal as an
from sck_amd import system as sy

ROW = [0.2, 0.0, -0.05, 0.1, 0.3, 0.4, 0.5]          # Jex2, Jex3, Jex4, Kan1, KanS, Hex, KDM


def test_build_state_is_the_nanodot_of_the_row():
    s = an.build_state(ROW, 3.0, 3, 30.0, shells=(1, 2, 4))
    r = sy.nanodot(Rd=3.0, L=3, Jex=1.0, Kan1=0.1, Kan2=0.0, KanS=0.3, Hex=0.4, KDM=0.5, gamma_deg=30.0,
                   shells=(1, 2, 4), J_shells={2: 0.2, 3: 0.0, 4: -0.05})
    assert s.J == r.J and np.array_equal(s.site_mask(), r.site_mask())
    assert np.allclose(s.h, r.h) and np.allclose(s.cubic_k1, r.cubic_k1) and s.cubic_k2 == r.cubic_k2 == 0.0
    assert len(s.crystal.bonds) == len(r.crystal.bonds)
    for i in range(len(s.crystal.bonds)):
        assert np.array_equal(s.bond_matrix(i), r.bond_matrix(i))
    with pytest.raises(ValueError, match="parameters"):
        an.build_state(ROW[:5], 3.0, 3, 30.0)


def test_active_shells_union():
    rows = [[0, 0, 0, 1, 1, 1, 1], [0.1, 0, 0, 1, 1, 1, 1], [0, 0, -0.2, 0, 0, 0, 0]]
    assert an.active_shells(rows) == [1, 2, 4]
    assert an.active_shells([[0] * 7]) == [1]


def test_reduce_obs_equals_direct_statistics():
    """Accumulators about arbitrary references give the replica means, the per-replica variances (Cv, chi) and
    the spreads that direct statistics of the series give."""
    rng = np.random.default_rng(4)
    nT, B, m, N = 3, 5, 400, 37
    temps = np.array([50.0, 20.0, 5.0])
    E = rng.normal(-80.0, 4.0, (nT, B, m))
    M = np.abs(rng.normal(0.6, 0.05, (nT, B, m)))
    Mz = rng.normal(0.1, 0.02, (nT, B, m))
    Q = rng.normal(0.0, 0.3, (nT, B, m))
    Er, Mr = E[..., 0] + 1.7, M[..., 0] - 0.2                    # references: any value works
    dE, dM = E - Er[..., None], M - Mr[..., None]
    obs = np.stack([dE.mean(-1), (dE ** 2).mean(-1), dM.mean(-1), (dM ** 2).mean(-1), Mz.mean(-1), Q.mean(-1),
                    Er, Mr], axis=1)
    r = an.reduce_obs(obs, temps, N)
    T = temps[:, None]
    assert np.allclose(r["E"], E.mean(-1).mean(1) / N)
    assert np.allclose(r["E_sd"], E.mean(-1).std(1) / N)
    assert np.allclose(r["M"], M.mean(-1).mean(1))
    assert np.allclose(r["Cv"], (E.var(-1) / (an.KB * T ** 2 * N)).mean(1), rtol=1e-9)
    assert np.allclose(r["chi"], (M.var(-1) / (an.KB * T)).mean(1), rtol=1e-9)
    assert np.allclose(r["Mz"], Mz.mean(-1).mean(1)) and np.allclose(r["Q"], Q.mean(-1).mean(1))
    with pytest.raises(ValueError):
        an.reduce_obs(obs[:, :7], temps, N)


def _config(path, **kw):
    cfg = dict(params=[ROW, [0.0] * 3 + ROW[3:]], names=["a", "b"], T_start=60.0, T_end=5.0, n_temps=3,
               sweeps=30, Rd=3.0, layers=3, angle=90.0, replicas=8, every=5, seed=7)
    cfg.update(kw)
    path.write_text(json.dumps(cfg))
    return path


def test_main_refuses_state_indices_out_of_range(tmp_path):
    with pytest.raises(SystemExit, match="state indices"):
        an.main(["--config", str(_config(tmp_path / "run.json")), "--states", "0,2", "--out", str(tmp_path / "o")])


# ----------------------------------------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_check_exact_counts_no_difference():
    print("[gpu] anneal check_exact", flush=True)
    assert an.check_exact(an.build_state(ROW, 3.0, 3, 90.0, shells=(1, 2, 4)), "cuda") == 0


@pytest.mark.gpu
def test_ladder_equals_the_same_chain_driven_eagerly():
    """Graph replays (10-sweep thermalization blocks, k sweeps plus the observation) and device accumulators
    give the accumulators of the same chain run with plain sweeps and observations."""
    from sck_amd.observables import topological_charge
    from sck_amd.sampler import Sampler
    print("[gpu] anneal ladder == eager chain", flush=True)
    s = an.build_state(ROW, 3.0, 3, 90.0, shells=(1, 2, 4))
    temps = [60.0, 25.0, 5.0]
    B, n_therm, n_meas, k, seed = 8, 23, 20, 4, 5
    r = an.anneal_state(s, B, temps, n_therm, n_meas, k, seed, "cuda", log=lambda m: None)
    A = Sampler(s, B, "cuda", seed)
    S = A.init_spins(seed=seed)
    n_act = float(s.n_active)
    m = n_meas // k
    for t, TK in enumerate(temps):
        A.set_temperature(an.KB * TK)
        A.sweep(S, n_therm)
        E0, M0 = A.observe(S)
        Er, Mr = E0, (M0.sum(1) / n_act).norm(dim=1)
        acc = torch.zeros(6, B, dtype=torch.float64, device=S.device)
        for _ in range(m):
            A.sweep(S, k)
            E, M = A.observe(S)
            Mv = M.sum(1) / n_act
            dE, dM = E - Er, Mv.norm(dim=1) - Mr
            acc += torch.stack([dE, dE * dE, dM, dM * dM, Mv[:, 2], topological_charge(S, s, s.L3 // 2)])
        got = r["obs"][t]
        assert np.allclose(got[:6], (acc / m).cpu().numpy(), rtol=1e-12, atol=1e-12), t
        assert np.array_equal(got[6], Er.cpu().numpy()) and np.allclose(got[7], Mr.cpu().numpy(), rtol=1e-14)
        assert np.array_equal(r["snaps"][t], S[0, 2, 0, s.L3 // 2].float().cpu().numpy()), t
    assert r["sweeps_per_chain"] == len(temps) * (n_therm + m * k)
    red = an.reduce_obs(r["obs"], temps, n_act)
    assert np.all(red["Cv"] >= 0) and np.all(red["chi"] >= 0) and np.all(np.abs(red["M"]) <= 1 + 1e-9)


@pytest.mark.gpu
def test_main_writes_one_file_per_state(tmp_path):
    print("[gpu] anneal main", flush=True)
    out = tmp_path / "out"
    an.main(["--config", str(_config(tmp_path / "run.json")), "--states", "1", "--out", str(out)])
    z = np.load(out / "state_b.npz")
    assert z["obs"].shape == (3, 8, 8) and z["M"].shape == (8, 3, 3) and z["snaps"].shape[0] == 3
    assert np.allclose(z["temperatures"], [60.0, 32.5, 5.0]) and int(z["n_active"]) > 0
    assert np.array_equal(z["params"], np.asarray([0.0] * 3 + ROW[3:], np.float32))


@pytest.mark.gpu
def test_anneal_ladder_matches_reference():
    """A nanodot cooling ladder (Rd = 2, two layers, 30 degree field, two temperatures) against the accumulators,
    magnetization vectors, middle-layer snapshots and reduced observables the reference's anneal_state and
    reduce_obs produced on its CPU path (SCK/anneal.py:143-234), from the same seed: fp32 decision tolerance."""
    from pathlib import Path

    from make_golden import ANNEAL_CASE as c
    from sck_amd import anneal as AN
    z = np.load(Path(__file__).resolve().parent / "golden" / "anneal_nanodot.npz")
    s = AN.build_state(c["params"], c["Rd"], c["L"], c["angle"], (1,))
    assert s.n_active == int(z["n_active"])
    r = AN.anneal_state(s, c["replicas"], c["temps_K"], c["n_therm"], c["n_meas"], c["every"], c["seed"],
                        device="cuda:0", log=lambda m: None)
    assert np.allclose(r["obs"], z["obs"], rtol=1e-4, atol=1e-4)
    assert np.allclose(r["mvec"], z["mvec"], atol=1e-5) and np.allclose(r["snaps"], z["snaps"], atol=1e-5)
    red = AN.reduce_obs(r["obs"], c["temps_K"], s.n_active)
    for k in ("E", "M", "Mz", "Cv", "chi", "Q"):
        assert np.allclose(red[k], z[f"reduced/{k}"], rtol=1e-4, atol=1e-4), k
