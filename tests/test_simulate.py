"""The run driver and its on-disk format (SCK/simulate.py, SCK/api.py).

CPU: shapes in three coordinate systems against brute force and the refusals of the shape checker, the
temperature ladder, replica blocks, the files `prepare` writes, `collect` on synthetic state blocks (Cv and chi
from the per-replica variances), and `check` reporting errors without raising.
GPU: a small two-state run end to end through `api.run` (files, shapes, physics of a field), two replica blocks
written by two `run` calls and combined by `collect`, and run-to-run determinism.
"""

import json
import math

import numpy as np
import pytest

from sck_amd import api
from sck_amd import simulate as SIM


def test_shapes_against_brute_force():
    _, m = SIM.active_sites("sc", [9, 9, 9], "spherical", "r <= 3")
    assert m.sum() == sum(1 for x in range(-4, 5) for y in range(-4, 5) for z in range(-4, 5) if x * x + y * y + z * z <= 9)
    _, m = SIM.active_sites("sc", [10, 10, 4], "cylindrical", "(rho <= 3.5) & (abs(z) <= 1)")
    xs = np.arange(10) - 4.5
    zs = np.arange(4) - 1.5
    want = sum(1 for x in xs for y in xs for z in zs if math.hypot(x, y) <= 3.5 and abs(z) <= 1)
    assert m.sum() == want
    _, m = SIM.active_sites("square", [6, 4], "cartesian", "x > 0")
    assert m.shape == (1, 1, 4, 6) and m.sum() == 12
    _, m = SIM.active_sites("bcc", [3, 3, 3])
    assert m.all() and m.shape == (2, 3, 3, 3)


@pytest.mark.parametrize("shape,msg", [("__import__('os')", "calls something other than"),
                                       ("(x > 0) and (y > 0)", "use & (and)"),
                                       ("q < 2", "name 'q' is not defined"),
                                       ("x < 'a'", "only numbers as constants"),
                                       ("x.real > 0", "may hold only numbers"),
                                       ("x > 100", "contains no site")])
def test_shape_checker_refusals(shape, msg):
    with pytest.raises(ValueError, match=msg.replace("(", r"\(").replace(")", r"\)")):
        SIM.active_sites("sc", [4, 4, 4], "cartesian", shape)


def test_ladder_and_replica_blocks():
    T = SIM.temperature_ladder(2.0, 0.2, 0.1)
    assert len(T) == 19 and T[0] == 2.0 and abs(T[-1] - 0.2) < 1e-12
    with pytest.raises(ValueError):
        SIM.temperature_ladder(1.0, 2.0, 0.1)
    assert SIM.replica_blocks(10, 3) == [(0, 4), (4, 7), (7, 10)]
    assert SIM.replica_blocks(2, 5) == [(0, 1), (1, 2)]


def _cfg(**kw):
    c = dict(model="heisenberg", crystal="sc", box=[4, 4, 4], periodic=[True] * 3,
             states=[{"name": "ferro", "J": {"1": 1.0}}, {"name": "field", "J": {"1": 1.0}, "H": 0.5}],
             T_max=2.0, T_min=1.0, T_step=0.5, thermalization=40, measurement=60, measure_every=5, replicas=8,
             config_every=2, seed=3)
    c.update(kw)
    return c


def test_prepare_writes_the_reference_files(tmp_path):
    run = SIM.prepare(_cfg(coords="spherical", shape="r <= 1.8"), tmp_path, log=lambda m: None)
    assert json.loads((tmp_path / "run.json").read_text())["n_sites"] == run["n_sites"]
    mask = np.load(tmp_path / "mask.npy")
    sites = np.load(tmp_path / "sites.npy")
    cf = np.load(tmp_path / "configurations.npy", mmap_mode="r")
    assert sites.shape == (mask.sum(), 3) and run["temperatures"] == [2.0, 1.5, 1.0]
    assert cf.shape == (2, 8, 2, int(mask.sum()), 3) and run["config_temperatures"] == [2.0, 1.0]


def test_collect_combines_blocks_and_derives_cv_chi(tmp_path):
    run = SIM.prepare(_cfg(), tmp_path, log=lambda m: None)
    rng = np.random.default_rng(0)
    nT, N = 3, run["n_sites"]
    full = {}
    for i in range(2):
        E = rng.normal(-100, 5, size=(8, nT))
        E2 = E ** 2 + rng.uniform(1, 3, size=(8, nT))
        Ma = rng.uniform(0.1, 0.9, size=(8, nT))
        M2 = Ma ** 2 + rng.uniform(0.001, 0.01, size=(8, nT))
        M = rng.normal(size=(8, nT, 3)).astype(np.float32)
        full[i] = (E, E2, Ma, M2)
        for part, (r0, r1) in enumerate([(0, 5), (5, 8)]):
            np.savez(tmp_path / f"state_{i:03d}_p{part:02d}.npz", r0=r0, r1=r1, M=M[r0:r1], E=E[r0:r1],
                     E2=E2[r0:r1], Mabs=Ma[r0:r1], M2=M2[r0:r1], config_temperatures=np.array([2.0, 1.0]))
    res = SIM.collect(tmp_path)
    T = np.array([2.0, 1.5, 1.0])
    for i in range(2):
        E, E2, Ma, M2 = full[i]
        assert np.allclose(res["Cv"][i], ((E2 - E ** 2) / (T ** 2 * N)).mean(axis=0))
        assert np.allclose(res["chi"][i], (N * (M2 - Ma ** 2) / T).mean(axis=0))
        assert np.allclose(res["E_site"][i], (E / N).mean(axis=0))
    assert np.load(tmp_path / "magnetization.npy").shape == (2, 8, 3, 3)
    (tmp_path / "state_001_p01.npz").unlink()
    with pytest.raises(ValueError, match="do not cover"):
        SIM.collect(tmp_path)


def test_check_reports_errors_without_raising():
    assert api.check(_cfg())["ok"]
    r = api.check(_cfg(model="potts", T_step=-1, replicas=0))
    assert not r["ok"] and len(r["errors"]) >= 3
    r = api.check(_cfg(crystal="sc", box=[4, 4]))
    assert not r["ok"]
    r = api.check(_cfg(states=[{"name": "a", "J": {"1": 1.0}}, {"name": "a", "J": {"1": 2.0}}]))
    assert any("distinct" in e for e in r["errors"])
    d = api.describe()
    assert "heisenberg" in d["models"] and "sc" in d["crystals"]


# ------------------------------------------------------------------------------------------------ GPU
def test_check_reports_the_size_of_a_valid_run():
    """For a valid configuration `check` also reports its sites, each state's colouring, the ladder and the
    budget (the keys of SCK/agent.py:117-128, read by the reference's nanodisk notebook): here a Heisenberg
    disc of radius 3.5 in an open 9 x 9 x 3 sc box, the surface counted site by site (fewer than six active
    neighbours), the budget from the settings."""
    cfg = dict(model="heisenberg", crystal="sc", box=[9, 9, 3], periodic=[False, False, False],
               coords="cylindrical", shape="rho <= 3.5", states=[dict(name="disc", J={1: 1.0}, K1=0.1)],
               T_max=2.0, T_min=0.5, T_step=0.5, thermalization=30, measurement=20, measure_every=5,
               replicas=8, config_every=2)
    r = api.check(cfg)
    assert r["ok"], r["errors"]
    xs = np.arange(9) - 4.0
    disc = {(x, y) for x in range(9) for y in range(9) if math.hypot(xs[x], xs[y]) <= 3.5}
    active = {(x, y, z) for (x, y) in disc for z in range(3)}
    nbrs = ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))
    on_surface = sum(1 for (x, y, z) in active if sum((x + a, y + b, z + c) in active for a, b, c in nbrs) < 6)
    assert r["sites"] == {"box": 243, "active": len(active), "surface": on_surface,
                          "bulk": len(active) - on_surface}
    assert [st["name"] for st in r["states"]] == ["disc"] and r["states"][0]["colours"] == 2
    assert r["states"][0]["periodic_used"] == [False, False, False]
    assert r["temperatures"] == {"count": 4, "first": 2.0, "last": 0.5}
    assert r["budget"] == {"chains": 8, "spin_updates": 8 * 4 * len(active) * 50,
                           "configurations_saved_per_chain": 2,
                           "output_gb": round(12.0 * 8 * 2 * len(active) / 1e9, 4)}
    assert not any(k in api.check(dict(cfg, replicas=0)) for k in ("sites", "budget"))
    d = api.describe()
    assert {"hamiltonian", "shape_functions", "shape_rules", "state_parameters", "defaults"} <= set(d)
    import sck_amd as sck                           # the reference's short interface: import sck; sck.check(...)
    assert sck.check is api.check and sck.run is api.run and sck.Results is api.Results
    assert "rho" in d["coordinates"]["cylindrical"] and "sqrt" in d["shape_functions"]


@pytest.mark.gpu
def test_run_end_to_end(tmp_path):
    res = api.run(_cfg(), output_dir=tmp_path, verbose=False)
    assert res.specific_heat.shape == (2, 3) and res.energy_per_site.shape == (2, 3)
    assert np.all(res.energy_per_site < 0)
    # a field orders the spins: |M| larger than without it, at every temperature
    assert np.all(res.magnetization_per_site[1] > res.magnetization_per_site[0])
    cf = res.configurations()
    norms = np.linalg.norm(cf, axis=-1)
    assert cf.shape == (2, 8, 2, 64, 3) and np.allclose(norms, 1.0, atol=1e-5)
    assert len(res.table()) == 6 and "ferro" in repr(res)


@pytest.mark.gpu
def test_replica_blocks_and_determinism(tmp_path):
    cfg = SIM.prepare(_cfg(), tmp_path / "a", log=lambda m: None)
    SIM.run(cfg, [0, 1], tmp_path / "a", "cuda:0", log=lambda m: None, replicas=(0, 3), part=0)
    SIM.run(cfg, [0, 1], tmp_path / "a", "cuda:0", log=lambda m: None, replicas=(3, 8), part=1)
    a = SIM.collect(tmp_path / "a")
    cfg = SIM.prepare(_cfg(), tmp_path / "b", log=lambda m: None)
    SIM.run(cfg, [0, 1], tmp_path / "b", "cuda:0", log=lambda m: None, replicas=(0, 3), part=0)
    SIM.run(cfg, [0, 1], tmp_path / "b", "cuda:0", log=lambda m: None, replicas=(3, 8), part=1)
    b = SIM.collect(tmp_path / "b")
    for k in ("energy", "Mabs", "Cv", "chi"):
        assert np.array_equal(a[k], b[k]), k
    assert a["energy"].shape == (2, 8, 3)


# ------------------------------------------------------------------ against fixtures written by the reference
from pathlib import Path  # noqa: E402

GOLDEN = Path(__file__).resolve().parent / "golden"


def _build_cases():
    from make_golden import BUILD_CASES
    return BUILD_CASES


@pytest.mark.parametrize("key", list(_build_cases()))
def test_build_system_matches_reference(key):
    """The state-keyed Hamiltonian front (J shells, bulk DMI, K1 with Ks on the surface, K2, field angles, masks,
    the open-box fallback) against the System the reference's build_system made for the same state
    (SCK/simulate.py:153-240)."""
    from sck_amd.hamiltonian import build_system
    model, crystal, box, per, theta, (coords, shape) = _build_cases()[key]
    z = np.load(GOLDEN / "build_system.npz")
    ref = {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(key + "/")}
    _, mask = SIM.active_sites(crystal, box, coords, shape)
    s = build_system(model, crystal, box, per, theta, mask, log=lambda m: None)
    assert s.n == int(ref["n"]) and list(s.box) == ref["box"].tolist()
    assert list(s.periodic) == ref["periodic"].tolist()               # includes the open-box fallback
    assert sorted(s.J) == ref["J_shells"].tolist() and [s.J[k] for k in sorted(s.J)] == ref["J_values"].tolist()
    assert sorted(s.extra) == ref["extra_bonds"].tolist()
    for i, E in zip(ref["extra_bonds"], ref["extra"]):
        assert np.allclose(s.extra[int(i)], E, rtol=0, atol=1e-15)
    assert np.allclose(np.asarray(s.h if s.h is not None else [], float), ref["h"], rtol=0, atol=1e-15)
    k1 = np.asarray(s.cubic_k1 if s.cubic_k1 is not None else np.zeros(0), np.float32)
    assert k1.shape == ref["k1"].shape and np.array_equal(k1, ref["k1"])
    assert s.cubic_k2 == float(ref["k2"]) and np.array_equal(s.site_mask(), ref["mask"])
    assert s.name == str(ref["name"])


def _run_case(key):
    from make_golden import RUN_CASES
    return RUN_CASES[key]


@pytest.mark.gpu
@pytest.mark.parametrize("key", ["run_ising", "run_heis"])
def test_run_folder_matches_reference(key, tmp_path):
    """prepare / run / collect on the GPU against the run folder the reference's driver wrote on its CPU path
    for the same configuration (SCK/simulate.py:432-530): the same run.json, sites, mask, configurations, state
    blocks and combined results. Ising chains are bit-exact, so their observables agree to float64 summation
    order; the Heisenberg run (DMI, K1 with Ks on a spherical surface, a tilted field, meV units) is compared at
    the fp32 decision tolerance."""
    cfg = _run_case(key)
    print(f"[gpu] reference run folder {key}", flush=True)
    z = np.load(GOLDEN / f"{key}.npz")
    run = SIM.prepare(cfg, tmp_path, log=lambda m: None)
    SIM.run(run, list(range(len(cfg["states"]))), tmp_path, device="cuda:0", log=lambda m: None)
    res = SIM.collect(tmp_path)
    assert json.loads((tmp_path / "run.json").read_text()) == json.loads(str(z["run_json"]))
    assert np.array_equal(np.load(tmp_path / "mask.npy"), z["mask"])
    assert np.allclose(np.load(tmp_path / "sites.npy"), z["sites"], rtol=0, atol=1e-6)
    exact = cfg["model"] == "ising"
    tol = dict(rtol=1e-9, atol=1e-9) if exact else dict(rtol=1e-4, atol=1e-4)
    cf = np.load(tmp_path / "configurations.npy")
    assert cf.shape == z["configurations"].shape
    assert np.array_equal(cf, z["configurations"]) if exact else np.allclose(cf, z["configurations"], atol=1e-5)
    for i in range(len(cfg["states"])):
        got = np.load(tmp_path / f"state_{i:03d}_p00.npz")
        for k in ("E", "E2", "Mabs", "M2", "M", "config_temperatures", "r0", "r1"):
            assert np.allclose(got[k], z[f"state_{i:03d}/{k}"], **tol), (i, k)
    for k in ("temperatures", "energy", "energy2", "Mabs", "M2", "Cv", "chi", "E_site", "M_site", "magnetization"):
        assert np.allclose(res[k], z[f"results/{k}"], **tol), k
    assert res["state_names"].tolist() == z["results/state_names"].tolist()
