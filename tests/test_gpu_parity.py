"""Parity of the HIP sweep path with the reference's trajectories and the CPU oracle (needs a HIP device).

Bar: Ising (int8, float32 storage, packed) and bitwise Ising bit-exact; XY/Heisenberg equal decisions except where
the oracle's margin |beta dE + ln u| < 1e-5, spin values within 1e-5; observables within 1e-9 relative.
"""

import numpy as np
import pytest
import torch

import oracle as O
from recipes import RECIPES
from sck_amd import crystals as cr
from sck_amd import system as sy
from sck_amd.sampler import Sampler

pytestmark = pytest.mark.gpu
GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden"
DEV = "cuda"


def _system(rc):
    s = eval(rc["system"], {"sy": sy, "cr": cr, "np": np})
    if rc.get("acceptance"):
        s.acceptance = rc["acceptance"]
    return s


def _sampler_from_recipe(rc, s):
    kw = dict(rc.get("sampler", {}))
    A = Sampler(s, rc["B"], device=DEV, seed=rc["seed"], **kw)
    A.set_temperature(torch.as_tensor(rc["T"], dtype=torch.float32))
    return A


# -------------------------------------------------------------------- against the reference's trajectories
@pytest.mark.parametrize("name", list(RECIPES))
def test_matches_reference_trajectory(name):
    rc = RECIPES[name]
    print(f"[gpu] reference trajectory {name}", flush=True)
    z = np.load(GOLDEN / f"{name}.npz")
    s = _system(rc)
    A = _sampler_from_recipe(rc, s)
    assert np.array_equal(A.beta.cpu().numpy(), z["beta"])
    if A.bitwise:
        assert np.array_equal(A.ptab.cpu().numpy(), z["ptab"])
    S = torch.as_tensor(z["S0"]).to(DEV).contiguous()
    assert S.dtype == A.init_spins(seed=rc["init_seed"]).dtype
    assert torch.equal(S, A.init_spins(seed=rc["init_seed"]))         # the reference's initialisation
    for k in range(rc["sweeps"]):
        mode = rc.get("mode", "sweep")
        if mode == "sync":
            A.sync_sweep(S, 1)
        elif mode == "reference":
            A.reference_sweep(S, 1)
        else:
            A.sweep(S, 1)
        got = S.cpu().numpy()
        if s.n == 1:
            assert np.array_equal(got, z["S"][k]), f"sweep {k}: {int((got != z['S'][k]).sum())} spins differ"
        else:
            assert np.allclose(got, z["S"][k], rtol=0, atol=1e-5), f"sweep {k}"
    E, M = A.energy(S).cpu().numpy(), A.magnetization(S).cpu().numpy()
    assert np.allclose(E, z["E"], rtol=1e-9, atol=1e-4)
    assert np.allclose(M, z["M"], rtol=1e-9, atol=1e-4)


# ----------------------------------------------------------------------------- parallel == sequential
SMALL = [lambda: sy.ising("square", 4), lambda: sy.xy("square", 4), lambda: sy.heisenberg("bcc", 2),
         lambda: sy.heisenberg("kagome", 3, J=-1.0), lambda: sy.nanodot(Rd=2.0, L=2),
         lambda: sy.System("Ising sc odd periodic", 1, cr.make("sc"), (5, 5, 5), (True,) * 3, {1: 1.0}),
         lambda: sy.System("Heisenberg sc 6x5x5 periodic", 3, cr.make("sc"), (6, 5, 5), (True,) * 3, {1: 1.0}),
         lambda: sy.heisenberg("fcc", 2),
         lambda: sy.System("XY sc 10x5x5 periodic", 2, cr.make("sc"), (10, 5, 5), (True,) * 3, {1: 1.0}),
         lambda: sy.nanodot(Rd=2.5, L=3, shells=(1, 2, 4), J_shells={2: 0.3, 4: -0.1}),
         lambda: sy.heisenberg("triangular", 24), lambda: sy.xy("square", 32), lambda: sy.ising("square", 32)]


@pytest.mark.parametrize("i", range(len(SMALL)))
def test_parallel_equals_sequential_chain(i):
    s = SMALL[i]()
    print(f"[gpu] parallel == sequential {s.name} {s.box}", flush=True)
    A, R = Sampler(s, 3, DEV, seed=11), Sampler(s, 3, DEV, seed=11)
    for X in (A, R):
        X.set_temperature(torch.tensor([0.7, 1.5, 3.0]))
    S1 = A.init_spins(seed=2)
    S2 = S1.clone()
    for _ in range(2):
        A.sweep(S1)
        R.reference_sweep(S2)
        assert torch.equal(S1, S2)
    assert not torch.equal(S1, A.init_spins(seed=2))
    assert int(A.state[1]) == 2 and int(R.state[1]) == 2


@pytest.mark.parametrize("i", [0, 2, 4, 5, 9, 12])
def test_resident_kernel_equals_class_launches(i):
    """Whole sweeps in LDS (one launch for n sweeps) equal the per-class launches spin for spin (Ising square 32
    with 64 threads: 8 sites per thread and class, read in two groups of 4 before their draws)."""
    s = SMALL[i]()
    print(f"[gpu] resident == class launches {s.name} {s.box}", flush=True)
    C = Sampler(s, 5, DEV, seed=4, resident=False)
    C.set_temperature(torch.tensor([0.5, 1.0, 2.0, 3.0, 5.0]))
    S0 = C.init_spins(seed=7)
    S2 = S0.clone()
    C.sweep(S2, 3)
    for rt in (None, 64, 1024):                       # the default and both ends of the thread-count range
        A = Sampler(s, 5, DEV, seed=4, resident_threads=rt)
        assert A.module.resident is not None
        A.set_temperature(torch.tensor([0.5, 1.0, 2.0, 3.0, 5.0]))
        S1 = S0.clone()
        A.sweep(S1, 3)
        assert torch.equal(S1, S2), rt
        assert int(A.state[1]) == 3


@pytest.mark.parametrize("make", [
    lambda: sy.System("Heis bcc J1J2 D h", 3, cr.make("bcc", (1, 2)), (4, 4, 6), (True,) * 3, {1: 1.0, 2: -0.2},
                      h=(0.0, 0.0, 0.05), uniaxial=(0.1, (0.0, 0.0, 1.0))),
    lambda: sy.xy("triangular", 12), lambda: sy.nanodot(Rd=3.0, L=3), lambda: sy.ising("square", 16)])
def test_replica_major_layout_equals_reference_layout(make):
    """Lanes over replicas ([n][site][B] storage) and lanes over x (reference layout) run the same update: equal
    spins; a reference-layout tensor handed to a replica-major sampler is converted and written back."""
    s = make()
    B = 128
    R = Sampler(s, B, DEV, seed=6, layout="replica", packed=False)
    F = Sampler(s, B, DEV, seed=6, layout="reference", resident=False, packed=False)
    assert R.replica_major and not F.replica_major
    T = torch.linspace(0.4, 3.0, B)
    for X in (R, F):
        X.set_temperature(T)
    S1, S2 = R.init_spins(seed=3), F.init_spins(seed=3)
    assert not S1.is_contiguous() and torch.equal(S1, S2)
    R.sweep(S1, 2)
    F.sweep(S2, 2)
    assert torch.equal(S1, S2)
    S3 = S2.clone()
    R.sweep(S3, 1)                                           # contiguous input: converted, then written back
    F.sweep(S2, 1)
    assert torch.equal(S3, S2)
    E1, M1 = R.observe(S1)
    E2, M2 = F.observe(S1.contiguous())
    assert torch.allclose(E1, E2, rtol=1e-12) and torch.allclose(M1, M2, rtol=1e-12)
    Rq = Sampler(s, B, DEV, seed=6, layout="replica", packed=False)
    Rq.set_temperature(T)
    S4 = Rq.init_spins(seed=3)
    Rq.reference_sweep(S4, 2)
    assert torch.equal(S4, S1.contiguous().view_as(S4)) or torch.equal(S4, S1)


def test_masked_sites_stay_zero_and_unit_norm_elsewhere():
    s = sy.nanodot(Rd=2.0, L=2)
    A = Sampler(s, 2, DEV, seed=3)
    S = A.init_spins(seed=0)
    A.sweep(S, 2)
    m = torch.as_tensor(s.site_mask(), device=DEV)
    norms = S.norm(dim=1)
    assert torch.all(norms[:, ~m] == 0)
    assert torch.allclose(norms[:, m], torch.ones_like(norms[:, m]), atol=1e-5)


@pytest.mark.parametrize("make", [
    lambda: sy.xy("triangular", 12),
    lambda: sy.heisenberg("bcc", 4),
    lambda: sy.nanodot(Rd=2.0, L=2),
    lambda: sy.ising("sc", 5),
])
def test_float64_storage_equals_float32(make):
    """spin_dtype=torch.float64, as the reference allows (its kernel loads any storage type as float32 and stores
    where(acc, t, s) cast back, SCK/generator.py:214, 310-319): sweeps, the sequential chain, the synchronous
    control, graph replays and the observables equal those of float32 storage exactly, and a sweep leaves every
    stored value a float32 value even where the move was rejected."""
    s = make()
    B = 6
    F = Sampler(s, B, DEV, seed=3, packed=False, spin_dtype=torch.float32)
    D = Sampler(s, B, DEV, seed=3, packed=False, spin_dtype=torch.float64)
    assert D.mode != F.mode and D.layout == "reference" and not D.use_tiled
    T = torch.linspace(0.6, 2.4, B)
    for X in (F, D):
        X.set_temperature(T)
    Sf = F.init_spins(seed=5)
    Sd = D.init_spins(seed=5)
    assert Sd.dtype == torch.float64 and torch.equal(Sd, Sf.double())
    F.sweep(Sf, 3)
    D.sweep(Sd, 3)
    assert torch.equal(Sd, Sf.double())
    Ef, Mf = F.observe(Sf)
    Ed, Md = D.observe(Sd)
    assert torch.equal(Ed, Ef) and torch.equal(Md, Mf)
    rf, rd = F.graph(Sf, 2), D.graph(Sd, 2)
    rf()
    rd()
    torch.cuda.synchronize()
    assert torch.equal(Sd, Sf.double())
    Rf, Rd = Sf.clone(), Sd.clone()
    F.reference_sweep(Rf, 1)
    D.reference_sweep(Rd, 1)
    assert torch.equal(Rd, Rf.double())
    F.sync_sweep(Sf, 1)
    D.sync_sweep(Sd, 1)
    assert torch.equal(Sd, Sf.double())
    if s.n > 1:
        # float64 values that float32 cannot hold: after one sweep every live site holds a float32 value
        X = torch.randn((B, s.n) + s.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(2))
        X = (X / X.norm(dim=1, keepdim=True)).to(DEV) * torch.as_tensor(s.site_mask(), device=DEV)
        assert not torch.equal(X, X.float().double())
        D.sweep(X, 1)
        assert torch.equal(X, X.float().double())


@pytest.mark.parametrize("make, B, kw", [
    (lambda: sy.ising("square", 4), 3, {}), (lambda: sy.xy("square", 4), 3, {}),
    (lambda: sy.heisenberg("bcc", 2), 3, {}), (lambda: sy.heisenberg("kagome", 3, J=-1.0), 3, {}),
    (lambda: sy.nanodot(Rd=2.0, L=2), 3, {}),
    (lambda: sy.xy("triangular", 128), 8, dict(tiled=True)),
    (lambda: sy.heisenberg("bcc", 6), 64, dict(layout="replica", resident=False)),
    (lambda: sy.ising("sc", 32), 32, dict(bitwise=True, independent=True)),
])
def test_zero_temperature_never_raises_energy(make, B, kw):
    """At beta -> infinity only dE <= 0 is accepted, so the energy cannot rise (the reference's
    tests/test_core.py:58-68 invariant, here on every kernel family: resident and class launches, the tiled
    sweep, the replica-major layout and the bitwise words)."""
    s = make()
    A = Sampler(s, B, DEV, seed=5, **kw)
    A.set_temperature(1e-6)
    S = A.init_spins(seed=1)
    E = A.energy(S)
    for _ in range(3):
        A.sweep(S)
        E2 = A.energy(S)
        assert torch.all(E2 <= E + 1e-4)
        E = E2


@pytest.mark.parametrize("make, base", [
    (lambda: sy.ising("square", 16), dict(bitwise=True)),
    (lambda: sy.ising("sc", 6), dict(bitwise=True, independent=True)),
    (lambda: sy.xy("square", 24), dict()),
])
def test_launch_options_change_no_result(make, base):
    """The reference's launch options (words per program, bitwise tile and warps, row compaction, max_block:
    SCK/sampler.py:40-131) choose launch geometry only; its tests check the stream is independent of them
    (tests/test_bitwise.py:84-97, 229-265). Every option gives the default's spins here too."""
    s = make()
    B = 96
    T = torch.linspace(1.2, 3.0, B // 32 if base.get("bitwise") else B)
    if base.get("bitwise"):
        T = T.repeat_interleave(32)
    runs = []
    for kw in ({}, dict(words_per_program=1), dict(words_per_program=4), dict(words_per_program=8),
               dict(bitwise_tile=64, bitwise_warps=2), dict(compact=False), dict(max_block=16)):
        A = Sampler(s, B, DEV, seed=11, **base, **kw)
        A.set_temperature(T)
        S = A.init_spins(seed=4)
        A.sweep(S, 3)
        runs.append(S.cpu())
    for i, r in enumerate(runs[1:], 1):
        assert torch.equal(r, runs[0]), i


def test_independent_bitwise_replicas_do_not_coalesce():
    """Ising square 8x8 at Tc, Barker, one word of 32 replicas, shared against independent draws (the reference's
    tests/test_bitwise.py:290-318). (a) From 32 identical configurations the shared draw keeps one configuration
    and the independent draws give at least 30 after 20 sweeps. (b) From random starts the shared draw collapses
    the word to a few configurations within 400 sweeps, while independent replicas share a configuration only
    where it is fully ordered."""
    s = sy.ising("square", 8)
    s.acceptance = "barker"
    tc = 2.0 / np.log(1.0 + np.sqrt(2.0))

    def sampler(independent):
        A = Sampler(s, 32, DEV, seed=5, bitwise=True, independent=independent)
        A.set_temperature(tc)
        return A

    def distinct(A, S):
        return len(np.unique(A.unpack(S)[:, 0].reshape(32, -1).cpu().numpy(), axis=0))

    res = {}
    for ind in (False, True):
        A = sampler(ind)
        x = A.unpack(A.init_spins(seed=2))
        S = A.pack(x[:1].expand_as(x).float())
        assert distinct(A, S) == 1
        A.sweep(S, 20)
        res[ind, "same"] = distinct(A, S)
        A = sampler(ind)
        S = A.init_spins(seed=2)
        assert distinct(A, S) == 32
        A.sweep(S, 400)
        res[ind, "random"] = distinct(A, S)
        x = A.unpack(S)[:, 0].reshape(32, -1).cpu().numpy()
        rows, counts = np.unique(x, axis=0, return_counts=True)
        res[ind, "repeated_disordered"] = int(sum(1 for r, c in zip(rows, counts) if c > 1 and abs(r.mean()) < 1))
    assert res[False, "same"] == 1 and res[True, "same"] >= 30, res
    assert res[False, "random"] <= 4 and res[True, "repeated_disordered"] == 0 and res[True, "random"] >= 28, res


def test_fully_masked_lattice_has_no_classes():
    """A lattice whose every site is masked has no classes (SCK/sampler.py:127-129 finds none): sweeps launch no
    class kernel, leave the spins zero and still advance the sweep counter, and E = M = 0."""
    s = sy.System("empty square", 1, cr.make("square"), (8, 8), (True, True), {1: 1.0},
                  mask=np.zeros((1, 1, 8, 8), bool))
    for kw in (dict(packed=False), dict(packed=False, spin_dtype=torch.float32)):
        A = Sampler(s, 4, DEV, seed=1, **kw)
        assert A.classes == [] and s.n_active == 0
        A.set_temperature(2.0)
        S = A.init_spins(seed=0)
        assert not torch.any(S != 0)
        A.sweep(S, 3)
        E, M = A.observe(S)
        assert not torch.any(S != 0) and int(A.state[1]) == 3
        assert torch.all(E == 0) and torch.all(M == 0)


def test_single_replica_batches_match_oracle():
    """B = 1 in every storage: int8 Ising bit-exact against the sequential oracle, a bitwise word holding one
    live replica (31 padding lanes) against the bitwise oracle, and Heisenberg decisions within the tolerance."""
    s = sy.ising("sc", 6)
    A = Sampler(s, 1, DEV, seed=4, packed=False)
    A.set_temperature(3.0)
    S = A.init_spins(seed=2)
    S0 = S.cpu().numpy()
    A.sweep(S, 4)
    ref, _ = O.chain(O.OracleSystem(s), S0.astype(np.float32), A.beta.cpu().numpy(), 4, 0, 4)
    assert np.array_equal(S.cpu().numpy().astype(np.float32), ref)
    W = Sampler(s, 1, DEV, seed=4, bitwise=True, independent=True)
    assert W.W == 1 and W.Bp == 32
    W.set_temperature(3.0)
    Sw = W.init_spins(seed=2)
    x0 = O.unpack(Sw.cpu().numpy(), 1, s.site_mask())[:, 0]
    W.sweep(Sw, 4)
    xw = O.bitwise_chain(O.OracleSystem(s), x0, W.ptab.cpu().numpy(), 4, 0, 4, independent=True)
    assert np.array_equal(O.unpack(Sw.cpu().numpy(), 1, s.site_mask())[:, 0], xw)
    h = sy.heisenberg("bcc", 4)
    H = Sampler(h, 1, DEV, seed=4)
    H.set_temperature(1.2)
    Sh = H.init_spins(seed=2)
    h0 = Sh.cpu().numpy()
    H.sweep(Sh, 1, classes=(0, 1))
    x, _, mg = O.chain(O.OracleSystem(h), h0, H.beta.cpu().numpy(), 4, 0, 1, class_range=(0, 1), margins=True)
    bad = np.abs(Sh.cpu().numpy() - x).max(axis=1) > 1e-5
    assert not np.any(bad & ~(np.abs(mg.reshape(bad.shape)) < 1e-5))


# ------------------------------------------------------------------------------ against the CPU oracle
def test_ising_c1_shape_bit_exact_against_oracle():
    """Ising square 64^2, int8, B = 512 (8 temperatures x 64 replicas in [1.5, 3.5]), seed 0: 30 sweeps on the GPU,
    replicas 0, 1, 200, 447, 511 followed by the sequential oracle with the same counters."""
    s = sy.ising("square", 64)
    T = np.repeat(np.linspace(1.5, 3.5, 8).astype(np.float32), 64)
    for resident in (True, False):
        A = Sampler(s, 512, DEV, seed=0, packed=False, resident=resident)
        assert (A.module.resident is not None) == resident      # without it: lanes over replicas
        A.set_temperature(torch.as_tensor(T))
        S = A.init_spins(seed=0)
        S0 = S.cpu().numpy()
        A.sweep(S, 30)
        reps = [0, 1, 200, 447, 511]
        ref, _ = O.chain(O.OracleSystem(s), S0[reps].astype(np.float32), A.beta.cpu().numpy(), 0, 0, 30, reps=reps)
        assert np.array_equal(S.cpu().numpy()[reps].astype(np.float32), ref), f"resident={resident}"


def test_ising_c1_config_1000_sweeps_bit_exact():
    """The whole C1 configuration: Ising square 64^2, int8, B = 512 (8 temperatures x 64 in [1.5, 3.5]), seed 0,
    1000 sweeps in one graph replay on the resident kernel; replicas 0 and 511 (the coldest and the hottest)
    equal the sequential oracle's 1000-sweep chains on the same counters, spin for spin."""
    s = sy.ising("square", 64)
    T = np.repeat(np.linspace(1.5, 3.5, 8).astype(np.float32), 64)
    A = Sampler(s, 512, DEV, seed=0, packed=False)
    A.set_temperature(torch.as_tensor(T))
    S = A.init_spins(seed=0)
    S0 = S.cpu().numpy()
    rp = A.graph(S, 1000)
    rp()
    torch.cuda.synchronize()
    assert int(A.state[1]) == 1000
    reps = [0, 511]
    ref, _ = O.chain(O.OracleSystem(s), S0[reps].astype(np.float32), A.beta.cpu().numpy(), 0, 0, 1000, reps=reps)
    assert np.array_equal(S.cpu().numpy()[reps].astype(np.float32), ref)


@pytest.mark.parametrize("workload", ["heisenberg", "xy"])
def test_full_size_vector_configs_match_oracle(workload):
    """The replica-major kernels at their real size: Heisenberg bcc 32^3, J1 = 1, J2 = -0.2, Du = 0.1, h = 0.05,
    B = 8192 (32 temperatures in [0.5, 3]), and XY triangular 2048^2, B = 2048 (16 temperatures in [0.7, 1.1]).
    For every class of sweep 0, replicas 0, B/2 + 1 and B - 1 are compared with the oracle from the same state:
    decisions equal except where |beta dE + ln u| < 1e-5; those replicas then continue from the oracle's state."""
    import bench
    s, kw, B, T, _ = bench._make(workload)
    print(f"[gpu] full-size {workload}: B={B}, {s.n_active} sites", flush=True)
    A = Sampler(s, B, DEV, seed=0)
    # Heisenberg bcc runs replica-major; the 2-D XY batch runs whole sweeps tiled, and per class (as here) the
    # reference-layout class kernels
    assert A.replica_major if workload == "heisenberg" else (A.use_tiled and A.layout == "reference")
    A.set_temperature(T)
    S = bench.device_spins(A, seed=0)
    reps = [0, B // 2 + 1, B - 1]
    os_ = O.OracleSystem(s)
    beta = A.beta.cpu().numpy()
    for c in range(len(A.classes)):
        x0 = S[reps].cpu().numpy()
        A.state[1] = 0
        A.sweep(S, 1, classes=(c, c + 1))
        ref, _, mg = O.chain(os_, x0, beta, 0, 0, 1, reps=reps, class_range=(c, c + 1), margins=True, threads=3)
        got = S[reps].cpu().numpy()
        bad = np.abs(got - ref).max(axis=1) > 1e-5
        near = np.abs(mg.reshape(bad.shape)) < 1e-5
        assert not np.any(bad & ~near), f"class {c}: {int((bad & ~near).sum())} decisions differ"
        assert int((np.abs(got - x0).max(axis=1) > 0).sum()) > 0            # the class did update sites
        S[reps] = torch.as_tensor(ref, device=DEV)


@pytest.mark.parametrize("acceptance", ["metropolis", "barker"])
def test_packed_matches_int8_and_oracle(acceptance):
    s = sy.System("Ising sc odd", 1, cr.make("sc"), (5, 5, 5), (True,) * 3, {1: 1.0}, h=(0.2,))
    s.acceptance = acceptance
    B = 40
    P = Sampler(s, B, DEV, seed=5)
    I8 = Sampler(s, B, DEV, seed=5, packed=False)
    F = Sampler(s, B, DEV, seed=5, packed=False, spin_dtype=torch.float32)
    assert P.packed and not I8.packed
    T = torch.linspace(1.5, 4.0, B)
    for X in (P, I8, F):
        X.set_temperature(T)
    Sp, S8, Sf = P.init_spins(seed=9), I8.init_spins(seed=9), F.init_spins(seed=9)
    S0 = S8.cpu().numpy()
    P.sweep(Sp, 3)
    I8.sweep(S8, 3)
    F.sweep(Sf, 3)
    assert torch.equal(P.unpack(Sp), S8)
    assert torch.equal(S8.float(), Sf)
    ref, _ = O.chain(O.OracleSystem(s), S0.astype(np.float32), I8.beta.cpu().numpy(), 5, 0, 3)
    assert np.array_equal(S8.cpu().numpy().astype(np.float32), ref)
    assert torch.allclose(P.energy(Sp), I8.energy(S8))
    Rp = Sampler(s, B, DEV, seed=5)
    Rp.set_temperature(T)
    Sr = Rp.init_spins(seed=9)
    Rp.reference_sweep(Sr, 3)
    assert torch.equal(Sr, Sp)


@pytest.mark.parametrize("independent", [False, True])
@pytest.mark.parametrize("acceptance", ["metropolis", "barker"])
def test_bitwise_against_oracle(independent, acceptance):
    s = sy.System("Ising sc 8 h", 1, cr.make("sc"), (8, 6, 4), (True, True, False), {1: 1.0}, h=(0.25,))
    s.acceptance = acceptance
    B = 32 * 6
    A = Sampler(s, B, DEV, seed=21, bitwise=True, independent=independent)
    A.set_temperature(torch.tensor([1.0, 2.0, 3.0, 4.0, 4.5, 6.0]).repeat_interleave(32))
    S = A.init_spins(seed=4)
    x = A.unpack(S).cpu().numpy()[:, 0]
    for k in range(3):
        A.sweep(S)
        x = O.bitwise_chain(O.OracleSystem(s), x, A.ptab.cpu().numpy(), 21, k, 1, independent=independent)
        assert np.array_equal(A.unpack(S).cpu().numpy()[:, 0], x), f"sweep {k}"
    R = Sampler(s, B, DEV, seed=21, bitwise=True, independent=independent)
    R.set_temperature(torch.tensor([1.0, 2.0, 3.0, 4.0, 4.5, 6.0]).repeat_interleave(32))
    S2 = R.init_spins(seed=4)
    R.reference_sweep(S2, 3)
    assert torch.equal(S2, S)


@pytest.mark.parametrize("h", [0.0, 0.3])
def test_bitwise_whole_row_waves_against_oracle(h):
    """Bitwise independent on a periodic square lattice whose class rows are whole waves (L = 128: 64 class
    slots per row), so a wave's row, its liveness and its words' table rows are uniform and a thread updates
    two words: every word against the oracle over 3 sweeps, with a field (which splits the entries by spin)
    and without, a different temperature per word."""
    s = sy.System("Ising square 128 h", 1, cr.make("square"), (128, 128), (True, True), {1: 1.0}, h=(h,))
    B = 32 * 4
    print(f"[gpu] bitwise whole-row waves h={h}", flush=True)
    A = Sampler(s, B, DEV, seed=13, bitwise=True, independent=True)
    A.set_temperature(torch.tensor([1.5, 2.27, 3.0, 5.0]).repeat_interleave(32))
    S = A.init_spins(seed=6)
    x = A.unpack(S).cpu().numpy()[:, 0]
    tab = A.ptab.cpu().numpy()
    for k in range(3):
        A.sweep(S)
        x = O.bitwise_chain(O.OracleSystem(s), x, tab, 13, k, 1, independent=True)
        assert np.array_equal(A.unpack(S).cpu().numpy()[:, 0], x), f"sweep {k}"


def test_bitwise_c2_shape_words_against_oracle():
    """Bitwise independent, sc 128^3 periodic, B = 1024 (32 words), T = 4.5: two sweeps on the GPU; words are
    independent chains, so the oracle follows words 0 and 31 over the whole lattice."""
    s = sy.ising("sc", 128)
    A = Sampler(s, 1024, DEV, seed=0, bitwise=True, independent=True)
    A.set_temperature(4.5)
    S = A.init_spins(seed=0)
    S0 = S.cpu().numpy()
    A.sweep(S, 2)
    got = S.cpu().numpy()
    os_ = O.OracleSystem(s)
    for wd in (0, 31):
        x = O.unpack(S0[wd:wd + 1], 32, s.site_mask())[:, 0]
        tab = A.ptab.cpu().numpy()[wd:wd + 1]
        x = O.bitwise_chain(os_, x, tab, 0, 0, 2, independent=True, word_offset=wd, threads=8)
        assert np.array_equal(O.unpack(got[wd:wd + 1], 32, s.site_mask())[:, 0], x), f"word {wd}"


VECTOR = [lambda: sy.heisenberg("bcc", 4), lambda: sy.xy("triangular", 12), lambda: sy.nanodot(Rd=3.0, L=3),
          lambda: sy.System("Heis bcc J1J2 D h", 3, cr.make("bcc", (1, 2)), (4, 4, 4), (True,) * 3, {1: 1.0, 2: -0.2},
                            h=(0.0, 0.0, 0.05), uniaxial=(0.1, (0.0, 0.0, 1.0)))]


@pytest.mark.parametrize("i", range(len(VECTOR)))
def test_vector_decisions_match_oracle_within_tolerance(i):
    """Per class launch from the same state: the GPU's accept decisions equal the oracle's except where the
    oracle's margin |beta dE + ln u| < 1e-5; accepted spins agree to 1e-5."""
    s = VECTOR[i]()
    B = 16
    A = Sampler(s, B, DEV, seed=7)
    A.set_temperature(torch.linspace(0.4, 3.0, B))
    S = A.init_spins(seed=5)
    os_ = O.OracleSystem(s)
    beta = A.beta.cpu().numpy()
    flips = 0
    for sweep in range(2):
        for c in range(len(A.classes)):
            x0 = S.cpu().numpy()
            A.state[1] = sweep
            A.sweep(S, 1, classes=(c, c + 1))
            ref, _, mg = O.chain(os_, x0, beta, 7, sweep, 1, class_range=(c, c + 1), margins=True)
            got = S.cpu().numpy()
            bad = np.abs(got - ref).max(axis=1) > 1e-5              # (B, b, L3, L2, L1)
            near = np.abs(mg.reshape(bad.shape)) < 1e-5
            assert not np.any(bad & ~near), f"sweep {sweep} class {c}: {int((bad & ~near).sum())} decisions differ"
            flips += int(bad.sum())
            S.copy_(torch.as_tensor(ref, device=DEV))               # continue from the oracle's state
    assert flips <= 2


def test_observables_match_oracle():
    for s in (sy.nanodot(Rd=3.0, L=3), sy.heisenberg("kagome", 6, J=-1.0), sy.ising("triangular", 6),
              sy.System("u", 3, cr.make("bcc", (1, 2)), (4, 4, 4), (True,) * 3, {1: 1.0, 2: -0.2},
                        h=(0.0, 0.0, 0.05), uniaxial=(0.1, (0.0, 0.0, 1.0)))):
        A = Sampler(s, 5, DEV, seed=1, packed=False)
        S = A.init_spins(seed=3)
        E, M = A.observe(S)
        Eo, Mo = O.observe(O.OracleSystem(s), S.cpu().numpy())
        assert np.allclose(E.cpu().numpy(), Eo, rtol=1e-12, atol=1e-9)
        assert np.allclose(M.cpu().numpy(), Mo, rtol=1e-12, atol=1e-9)


def test_set_seed_restarts_the_stream_in_place():
    """set_seed(seed, sweep) rewrites the state pair the kernels read (SCK/sampler.py:148-150): the chain then
    continues exactly as a Sampler built with that seed whose sweep counter stands at `sweep`, and a captured
    graph sees the new seed, as it sees beta."""
    s = sy.ising("square", 16)
    A = Sampler(s, 64, DEV, seed=1, packed=False)
    Bs = Sampler(s, 64, DEV, seed=9, packed=False)
    for X in (A, Bs):
        X.set_temperature(2.2)
    S1 = A.init_spins(seed=4)
    S2 = S1.clone()
    rep = A.graph(S1, 2)
    A.set_seed(9, 5)
    Bs.state[1] = 5
    rep()
    Bs.sweep(S2, 2)
    torch.cuda.synchronize()
    assert torch.equal(S1, S2)
    assert int(A.state[0]) == 9 and int(A.state[1]) == 7


def test_graph_replay_equals_eager_sweeps_and_reads_beta_by_address():
    s = sy.heisenberg("bcc", 4)
    A, Bs = Sampler(s, 8, DEV, seed=2), Sampler(s, 8, DEV, seed=2)
    for X in (A, Bs):
        X.set_temperature(1.2)
    S1 = A.init_spins(seed=1)
    S2 = S1.clone()
    rep = A.graph(S1, 3)
    rep()
    Bs.sweep(S2, 3)
    torch.cuda.synchronize()
    assert torch.equal(S1, S2) and int(A.state[1]) == 3
    A.set_temperature(0.5)                                     # in place: the graph sees it
    Bs.set_temperature(0.5)
    for _ in range(3):
        rep()
        Bs.sweep(S2, 3)
        torch.cuda.synchronize()
        assert torch.equal(S1, S2)
    assert int(A.state[1]) == 12


@pytest.mark.parametrize("make, B, blk, chunks", [
    (lambda: sy.System("Heis bcc J1J2 D h", 3, cr.make("bcc", (1, 2)), (4, 4, 4), (True,) * 3, {1: 1.0, 2: -0.2},
                       h=(0.0, 0.0, 0.05), uniaxial=(0.1, (0.0, 0.0, 1.0))), 256, 64, None),
    (lambda: sy.xy("triangular", 12), 512, 128, 2)])
def test_replica_blocks_equal_unblocked_sweeps(make, B, blk, chunks):
    """Replica blocks (each block through all n sweeps, then the next) leave every replica's chain unchanged:
    whole sweeps, a class range and graph replays give the spins of the unblocked sampler."""
    s = make()
    print(f"[gpu] replica blocks {s.name} B={B} block={blk}", flush=True)
    P = Sampler(s, B, DEV, seed=9, layout="replica", replica_block=blk, chunks=chunks)
    Q = Sampler(s, B, DEV, seed=9, layout="replica", replica_block=0, chunks=chunks)
    assert P.module.rep_blocks == B // blk and Q.module.rep_blocks == 1
    T = torch.linspace(0.3, 3.0, B)
    for X in (P, Q):
        X.set_temperature(T)
    S1, S2 = P.init_spins(seed=4), Q.init_spins(seed=4)
    P.sweep(S1, 3)
    Q.sweep(S2, 3)
    assert torch.equal(S1, S2)
    P.sweep(S1, 2, classes=(1, 3))
    Q.sweep(S2, 2, classes=(1, 3))
    assert torch.equal(S1, S2)
    rep = P.graph(S1, 2)
    for _ in range(2):
        rep()
        Q.sweep(S2, 2)
        torch.cuda.synchronize()
        assert torch.equal(S1, S2)
    ms, nl = P.sweep_timed(S1, 1)
    Q.sweep(S2, 1)
    assert nl == len(P.classes) * (B // blk) and torch.equal(S1, S2)
    assert int(P.state[1]) == int(Q.state[1]) == 10
    # an observing graph runs each block's observation right after that block's sweeps: every replay gives
    # the spins of the unblocked sampler, and observables equal to an eager observation of the same state
    # (and to the unblocked module's, up to float64 summation order)
    assert P.module.obs_blocks == B // blk
    obs = P.graph(S1, 2, observe=True)
    for _ in range(3):
        obs()
        Q.sweep(S2, 2)
        torch.cuda.synchronize()
        assert torch.equal(S1, S2)
        Eg, Mg = [x.clone() for x in P.observed()]
        Ee, Me = P.observe(S1)
        assert torch.equal(Eg, Ee) and torch.equal(Mg, Me)
        Eq, Mq = Q.observe(S2)
        assert torch.allclose(Ee, Eq, rtol=1e-12, atol=1e-9) and torch.allclose(Me, Mq, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("resident", [False, True])
def test_timed_sweeps_equal_plain_sweeps(resident):
    """sck_sweep_timed runs the same sweeps (spins and counter) and counts one event pair per sweep-kernel
    launch: one per class per sweep, or one resident launch for all n sweeps."""
    s = sy.ising("square", 16)
    print(f"[gpu] timed sweeps resident={resident}", flush=True)
    A, Bs = (Sampler(s, 64, DEV, seed=3, packed=False, resident=resident) for _ in range(2))
    assert (A.module.resident is not None and A.use_resident) == resident
    for X in (A, Bs):
        X.set_temperature(torch.linspace(1.0, 4.0, 64))
    S1 = A.init_spins(seed=2)
    S2 = S1.clone()
    ms, nl = A.sweep_timed(S1, 4)
    Bs.sweep(S2, 4)
    torch.cuda.synchronize()
    assert torch.equal(S1, S2) and int(A.state[1]) == 4
    assert nl == (1 if resident else 4 * len(A.classes))
    assert ms > 0.0
    with pytest.raises(ValueError):
        A.sweep_timed(S1, 0)


def test_sync_sweep_matches_oracle():
    s = sy.xy("square", 8)
    A = Sampler(s, 4, DEV, seed=3)
    A.set_temperature(0.9)
    S = A.init_spins(seed=1)
    x = O.sync(O.OracleSystem(s), S.cpu().numpy(), A.beta.cpu().numpy(), 3, 0, 1)
    A.sync_sweep(S, 1)
    assert np.allclose(S.cpu().numpy(), x, atol=1e-5)


def test_bitwise_needs_one_temperature_per_word():
    A = Sampler(sy.ising("square", 4), 64, DEV, bitwise=True)
    A.set_temperature(torch.tensor([2.0] * 32 + [3.0] * 32))
    with pytest.raises(ValueError, match="word"):
        A.set_temperature(torch.linspace(2.0, 3.0, 64))


@pytest.mark.parametrize("name, want", [("ising", (-2.000000, -2.000000)), ("xy", (-1.812110, -2.105182)),
                                        ("heisenberg", (-1.086448, -1.125840))])
def test_reference_quickstart_on_the_gpu(name, want):
    """The reference's quickstart end to end on the GPU Sampler (examples/quickstart.py:26-40, expected output
    examples/quickstart.expected.txt:1-3): B = 2, seed 11, T = 1.5, initial spins from seed 2; the coloured
    sweep equals the sequential chain after 3 sweeps, and the energy per site after 20 sweeps prints as the
    reference's: exactly for Ising, and for XY and Heisenberg within 2e-6 (the reference allows their last
    digits to differ on a GPU, whose float32 cos, sin and exp are not numpy's)."""
    s = {"ising": lambda: sy.ising("square", 8), "xy": lambda: sy.xy("sc", 4),
         "heisenberg": lambda: sy.heisenberg("triangular", 6)}[name]()
    print(f"[gpu] quickstart {s.name}", flush=True)
    A, R = Sampler(s, 2, DEV, seed=11), Sampler(s, 2, DEV, seed=11)
    A.set_temperature(1.5)
    R.set_temperature(1.5)
    S1 = A.init_spins(seed=2)
    S2 = S1.clone()
    A.sweep(S1, 3)
    R.reference_sweep(S2, 3)
    assert torch.equal(S1, S2)
    A.sweep(S1, 17)
    e = (A.energy(S1) / s.n_active).cpu().numpy()
    if s.n == 1:
        assert (round(float(e[0]), 6), round(float(e[1]), 6)) == want
    else:
        assert np.allclose(e, want, rtol=0, atol=2e-6), e
