"""Batched uneven systems (nanodisks with surface-normal anisotropy and DMI, uneven replica counts).

CPU: the surface normals against a brute-force loop over sites, the DMI matrix against the textbook form
D r . (S_i x S_j), the batch host tables (compacted class lists cover every active site of every class once;
every schedule's work list covers every (system, replica, class) range once; the XCD interleave keeps a
chain's tiles on one XCD), and the structure check.
The oracle's surface-anisotropy terms are pinned by mathematics: the energy of uniform states in closed form
and the single-spin distribution by quadrature over the sphere.
GPU: a batch equals independent Samplers of its systems bit for bit under every schedule the autotuner can
pick, for Heisenberg (uniform and cone proposals) and Ising; the batch's parallel sweep equals its own
sequential chain; batch observables equal the single-system kernels; graph replays equal eager sweeps; a
nanodisk's decisions match the oracle's; the sweep samples the single-spin distribution by quadrature.
"""

import itertools
import json
import math
from dataclasses import asdict

import numpy as np
import pytest
import torch

import oracle as O
from sck_amd import crystals as cr
from sck_amd import system as sy
from sck_amd.batch import XCDS, BatchLayout, BatchSampler, Schedule, candidate_schedules, signature
from sck_amd.hamiltonian import dmi, nanodisk, surface_normals

DEV = "cuda"


def _disks(nd=5, seed=0, rmax=6.0, tmax=4):
    rng = np.random.default_rng(seed)
    return [nanodisk(float(rng.uniform(2.5, rmax)), int(rng.integers(2, tmax + 1))) for _ in range(nd)]


# ------------------------------------------------------------------------------------------------ CPU
def test_surface_normals_match_brute_force():
    s = nanodisk(4.3, 3)
    c = s.crystal
    m = s.site_mask()
    got = surface_normals(c, s.box, s.periodic, m)
    L1, L2, L3 = s.box
    for z, y, x in itertools.product(range(L3), range(L2), range(L1)):
        if not m[0, z, y, x]:
            assert not got[0, z, y, x].any()
            continue
        acc = np.zeros(3)
        first = None
        for bd in c.bonds:
            if bd.shell != 1:
                continue
            xx, yy, zz = x + bd.delta[0], y + bd.delta[1], z + bd.delta[2]
            inside = 0 <= xx < L1 and 0 <= yy < L2 and 0 <= zz < L3
            if inside and m[0, zz, yy, xx]:
                continue
            r = np.array(bd.delta, float)
            acc += r
            first = r if first is None else first
        if first is None:
            assert not got[0, z, y, x].any()
            continue
        want = acc / np.linalg.norm(acc) if np.linalg.norm(acc) > 1e-9 else first / np.linalg.norm(first)
        assert np.allclose(got[0, z, y, x], want, atol=1e-12), (x, y, z)
    # the anisotropy sits on exactly the surface sites, along their normals
    ku = s.site_ku.reshape(m.shape)
    assert np.all((ku > 0) == (np.linalg.norm(got, axis=-1) > 0))
    assert np.allclose(ku[ku > 0], 0.3)


def test_dmi_is_the_textbook_cross_product_form():
    c = cr.make("sc")
    D = 0.1
    A = dmi(c, D)
    rng = np.random.default_rng(0)
    assert len(A) == 4                                        # in-plane first-shell bonds only
    for i, M in A.items():
        r = np.array(c.bonds[i].delta, float)
        assert r[2] == 0
        si, sj = rng.normal(size=3), rng.normal(size=3)
        assert math.isclose(si @ M @ sj, D * r @ np.cross(si, sj), rel_tol=1e-12, abs_tol=1e-15)
        assert np.allclose(M, -M.T)


def test_nanodisk_shapes():
    s = nanodisk(8, 4)
    assert s.box == (17, 17, 4) and s.n_active == 4 * int(sum(1 for x in range(-8, 9) for y in range(-8, 9)
                                                              if x * x + y * y <= 64))
    big = nanodisk(40, 12)
    assert big.box == (81, 81, 12)


def test_oracle_energy_of_uniform_nanodisk_states_is_the_closed_form():
    """All spins along u: DMI vanishes (S x S = 0), exchange gives -J1 per active bond, and the surface term
    -Ks (n . u)^2 per surface site."""
    s = nanodisk(5.2, 3)
    m = s.site_mask()
    bonds = int((m[0, :, :, :-1] & m[0, :, :, 1:]).sum()) + int((m[0, :, :-1, :] & m[0, :, 1:, :]).sum()) \
        + int((m[0, :-1] & m[0, 1:]).sum())
    ku = s.site_ku.reshape(-1)
    ax = s.site_axis.reshape(-1, 3).astype(np.float64)
    for u in (np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0]), np.array([0.6, 0.0, 0.8])):
        x = np.zeros((1, 3) + s.shape, np.float32)
        for k in range(3):
            x[0, k] = u[k] * m
        E, M = O.observe(O.OracleSystem(s), x)
        uu = x[0].reshape(3, -1).T.astype(np.float64)
        want = -1.0 * bonds - float((ku * ((ax * uu).sum(1)) ** 2).sum())
        assert abs(E[0] - want) < 1e-4 * abs(want), (u, E[0], want)
        assert np.allclose(M[0, 0], u * m.sum(), rtol=1e-6)


def _site_anisotropy_spin():
    ax = np.array([0.6, 0.0, 0.8])
    s = sy.System("single spin, site anisotropy", 3, cr.make("sc"), (1, 1, 1), (False,) * 3, {1: 1.0},
                  h=(0.0, 0.0, 0.4), site_ku=np.full((1, 1, 1, 1), 0.9, np.float32),
                  site_axis=np.broadcast_to(ax.astype(np.float32), (1, 1, 1, 1, 3)).copy())
    return s, ax, 0.9, 0.4, 0.5


def _sphere_average(ax, Ks, h, T):
    """(<s_z>, <(s.a)^2>) under exp(beta (h s_z + Ks (s.a)^2)) by quadrature over the sphere."""
    th = np.linspace(0, np.pi, 1201)
    ph = np.linspace(0, 2 * np.pi, 1201)
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    x, y, z = np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)
    dot = ax[0] * x + ax[1] * y + ax[2] * z
    wgt = np.exp((h * z + Ks * dot * dot) / T) * np.sin(TH)
    Z = np.trapezoid(np.trapezoid(wgt, ph, axis=1), th)
    return (np.trapezoid(np.trapezoid(wgt * z, ph, axis=1), th) / Z,
            np.trapezoid(np.trapezoid(wgt * dot * dot, ph, axis=1), th) / Z)


def test_oracle_site_anisotropy_single_spin_matches_quadrature():
    """Pins the oracle's per-site anisotropy term (sign, axis) by the exact single-spin distribution."""
    s, ax, Ks, h, T = _site_anisotropy_spin()
    B = 2048
    rng = np.random.default_rng(0)
    x = rng.normal(size=(B, 3, 1, 1, 1, 1)).astype(np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    os_ = O.OracleSystem(s)
    beta = np.full(B, 1.0 / T, np.float32)
    x, _ = O.chain(os_, x, beta, 4, 0, 200, threads=8)
    sz, sa = [], []
    for i in range(20):
        x, _ = O.chain(os_, x, beta, 4, 200 + 5 * i, 5, threads=8)
        v = x[:, :, 0, 0, 0, 0].astype(np.float64)
        sz.append(v[:, 2].mean())
        sa.append(((v @ ax) ** 2).mean())
    ez, ea = _sphere_average(ax, Ks, h, T)
    for series, exact in ((sz, ez), (sa, ea)):
        assert abs(np.mean(series) - exact) < 3 * np.std(series) / math.sqrt(len(series)) + 2e-3, exact


def _layout(systems, reps, seeds=None):
    from sck_amd import hipgen as G
    classes = systems[0].classes()
    seeds = seeds or list(range(len(systems)))
    return BatchLayout(systems, reps, seeds, classes, systems[0].n, systems[0].b), G


def test_batch_lists_cover_every_active_class_site_once():
    systems = _disks(4, seed=1)
    lay, G = _layout(systems, [2, 3, 1, 4])
    mask_bits = (1 << G.COORD_BITS) - 1
    for i, s in enumerate(systems):
        cm, m = s.colour_map(), s.site_mask()
        for k, (c, a) in enumerate(lay.classes):
            lo, cnt = lay.desc[i, 8 + 2 * k], lay.desc[i, 9 + 2 * k]
            p = lay.lists[lo:lo + cnt]
            x, y, z = p & mask_bits, (p >> G.COORD_BITS) & mask_bits, p >> (2 * G.COORD_BITS)
            want = np.argwhere((cm[a] == c) & m[a])                 # (z, y, x), raster order
            assert np.array_equal(np.stack([z, y, x], 1), want)
        assert tuple(lay.desc[i, :8]) == (s.L1, s.L2, s.L3, lay.reps[i], lay.rep_off[i], lay.seeds[i],
                                          lay.spin_off[i], lay.site_off[i])
    # the header and descriptor words sit where the kernels read them
    words = lay.aux.view(np.int64)
    assert words[0] == len(systems) and words[2] == len(lay.classes)
    assert np.array_equal(words[4:4 + lay.desc.size], lay.desc.reshape(-1))
    off = int(words[1])
    assert np.array_equal(lay.aux.view(np.uint32)[off:off + lay.lists.size], lay.lists)


@pytest.mark.parametrize("sch", candidate_schedules(), ids=lambda s: s.label())
def test_every_schedule_covers_every_range_once(sch):
    systems = _disks(5, seed=2)
    reps = [3, 1, 4, 2, 5]
    lay, G = _layout(systems, reps)
    if sch.kind == "chain":
        t = lay.chain_tiles(sch)
        real = t[t[:, 3] >= 0]
        assert sorted(map(tuple, real[:, :2])) == sorted(map(tuple, lay.chains))
        if sch.order == "xcd":
            q = lay.chain_order("xcd")
            for p, row in enumerate(t):
                if row[3] >= 0:
                    assert (int(row[0]), int(row[1])) in q[p % XCDS]
        return
    for k in range(len(lay.classes)):
        t = lay.class_tiles(k, sch)
        assert np.all(t[:, 3] - t[:, 2] <= sch.tile)
        cover = {}
        for sysi, rep, lo, hi in t:
            if hi > lo:
                cover.setdefault((int(sysi), int(rep)), []).append((int(lo), int(hi)))
        for i in range(len(systems)):
            lo0, cnt = int(lay.desc[i, 8 + 2 * k]), int(lay.desc[i, 9 + 2 * k])
            for r in range(reps[i]):
                segs = sorted(cover.pop((i, r)))
                assert segs[0][0] == lo0 and segs[-1][1] == lo0 + cnt
                assert all(a[1] == b[0] for a, b in zip(segs, segs[1:]))
        assert not cover
        if sch.order == "xcd":
            # tiles of one chain share a position residue mod 8: one XCD under any grid that is a multiple of 8
            res = {}
            for p, (sysi, rep, lo, hi) in enumerate(t):
                if hi > lo:
                    res.setdefault((int(sysi), int(rep)), set()).add(p % XCDS)
            assert all(len(v) == 1 for v in res.values())


def test_structure_mismatch_is_rejected():
    a = nanodisk(3, 2)
    b = nanodisk(3, 2, D=0.2)
    assert signature(a) != signature(b)
    assert signature(a) == signature(nanodisk(5.5, 4))


def test_batch_source_compiles_for_gfx950():
    from sck_amd import _lib as L
    from sck_amd import hipgen as G
    for systems, mode, prop in ((_disks(3), G.MODE_F32, "uniform"), (_disks(3), G.MODE_F32, "cone"),
                                ([sy.ising("sc", 4), sy.ising("sc", 6)], G.MODE_INT8, "uniform")):
        m = G.generate_batch(systems, mode, prop)
        for k in m.kernels:
            assert f" {k}(sck_args G)" in m.src or f" {k}(sck_args A)" in m.src
        L.precompile(m.src)


# ------------------------------------------------------------------------------------------------ GPU
def _individual(systems, reps, T, S0_views, n, seed0, **kw):
    from sck_amd.sampler import Sampler
    out = []
    for i, s in enumerate(systems):
        A = Sampler(s, reps[i], DEV, seed=seed0 + i, **kw)
        A.set_temperature(torch.as_tensor(T[i]))
        X = S0_views[i].clone()
        A.sweep(X, n)
        out.append(X)
    return out


@pytest.mark.gpu
def test_batch_equals_individual_samplers_under_every_schedule():
    systems = _disks(5, seed=3)
    reps = [3, 4, 5, 2, 4]
    T = [np.linspace(0.3, 2.0, r).astype(np.float32) for r in reps]
    Bt = BatchSampler(systems, reps, DEV, seed=11, schedule=Schedule())
    Bt.set_temperature(T)
    S0 = Bt.init_spins(seed=4)
    ref = _individual(systems, reps, T, Bt.views(S0), 3, 11)
    ref_nr = _individual(systems, reps, T, Bt.views(S0), 3, 11, resident=False)
    for a, b in zip(ref, ref_nr):
        assert torch.equal(a, b)
    for sch in candidate_schedules():
        print("schedule", sch.label(), flush=True)
        Bt.use(sch)
        Bt.state[1] = 0
        S = S0.clone()
        Bt.sweep(S, 3)
        for i, v in enumerate(Bt.views(S)):
            assert torch.equal(v, ref[i]), (sch.label(), i)
        assert int(Bt.state[1]) == 3


@pytest.mark.gpu
def test_batch_cone_proposal_and_ising_equal_individual_samplers():
    systems = _disks(3, seed=4)
    reps = [2, 3, 2]
    T = [np.full(r, 0.7, np.float32) for r in reps]
    widths = np.linspace(0.3, 2.5, sum(reps)).astype(np.float32)
    Bt = BatchSampler(systems, reps, DEV, seed=5, proposal="cone", width=widths, schedule=Schedule("chain", 256, 0, "xcd"))
    Bt.set_temperature(T)
    S0 = Bt.init_spins(seed=1)
    S = S0.clone()
    Bt.sweep(S, 2)
    from sck_amd.sampler import Sampler
    off = 0
    for i, s in enumerate(systems):
        A = Sampler(s, reps[i], DEV, seed=5 + i, proposal="cone", width=widths[off:off + reps[i]])
        A.set_temperature(torch.as_tensor(T[i]))
        X = Bt.views(S0)[i].clone()
        A.sweep(X, 2)
        assert torch.equal(X, Bt.views(S)[i]), i
        off += reps[i]
    # Ising: periodic sc boxes of different sizes, int8 spins and per-replica acceptance tables
    isys = [sy.ising("sc", 4), sy.ising("sc", 6), sy.ising("sc", 8)]
    ireps = [2, 1, 3]
    iT = [np.linspace(3.0, 6.0, r).astype(np.float32) for r in ireps]
    for sch in (Schedule(), Schedule("chain", 1024, 0, "largest"), Schedule("class", 64, 64, "system", 2048)):
        Bi = BatchSampler(isys, ireps, DEV, seed=9, schedule=sch)
        Bi.set_temperature(iT)
        I0 = Bi.init_spins(seed=2)
        I = I0.clone()
        Bi.sweep(I, 4)
        ref = _individual(isys, ireps, iT, Bi.views(I0), 4, 9)
        for i, v in enumerate(Bi.views(I)):
            assert torch.equal(v, ref[i]), (sch.label(), i)


@pytest.mark.gpu
def test_batch_parallel_sweep_equals_its_sequential_chain_and_graph_replay():
    systems = _disks(4, seed=5)
    reps = [2, 2, 3, 1]
    Bt = BatchSampler(systems, reps, DEV, seed=3, schedule=Schedule("class", 128, 512, "xcd", 2048))
    Bt.set_temperature(0.8)
    S0 = Bt.init_spins(seed=6)
    S1, S2, S3 = S0.clone(), S0.clone(), S0.clone()
    Bt.sweep(S1, 2)
    Bt.state[1] = 0
    Bt.reference_sweep(S2, 2)
    assert torch.equal(S1, S2)
    Bt.state[1] = 0
    rep = Bt.graph(S3, 1)
    rep()
    rep()
    torch.cuda.synchronize()
    assert torch.equal(S1, S3) and int(Bt.state[1]) == 2
    # timed sweeps: the same chain, one event pair per class launch
    Bt.state[1] = 0
    S4 = S0.clone()
    ms, nl = Bt.sweep_timed(S4, 2)
    assert torch.equal(S1, S4) and int(Bt.state[1]) == 2
    assert nl == 2 * len(Bt.classes) and ms > 0.0


@pytest.mark.gpu
def test_batch_observables_equal_single_system_kernels():
    from sck_amd.sampler import Sampler
    systems = _disks(3, seed=6)
    reps = [2, 3, 2]
    Bt = BatchSampler(systems, reps, DEV, seed=1, schedule=Schedule())
    S = Bt.init_spins(seed=8)
    E, M = Bt.observe(S)
    for i, (s, e, m) in enumerate(zip(systems, Bt.split(E), Bt.split(M))):
        A = Sampler(s, reps[i], DEV, seed=1)
        Ei, Mi = A.observe(Bt.views(S)[i].clone())
        assert torch.allclose(e, Ei, rtol=1e-12, atol=1e-9) and torch.allclose(m, Mi, rtol=1e-12, atol=1e-9)
        Eo, Mo = O.observe(O.OracleSystem(s), Bt.views(S)[i].cpu().numpy())
        assert np.allclose(e.cpu().numpy(), Eo, rtol=1e-12, atol=1e-9)


@pytest.mark.gpu
def test_autotuned_batch_keeps_the_chain():
    systems = _disks(6, seed=7)
    reps = 3
    Bt = BatchSampler(systems, reps, DEV, seed=2)            # schedule="auto"
    Bt.set_temperature(1.1)
    S0 = Bt.init_spins(seed=3)
    S = S0.clone()
    Bt.sweep(S, 2)                                           # tunes on a copy first
    assert Bt.tuning and Bt.schedule is not None and int(Bt.state[1]) == 2
    tuned = json.loads(Bt.schedule_json())                   # the chosen schedule and the timings behind it
    assert tuned["schedule"] == asdict(Bt.schedule) and tuned["tuning"] == json.loads(json.dumps(Bt.tuning))
    ref = _individual(systems, [reps] * 6, [np.full(3, 1.1, np.float32)] * 6, Bt.views(S0), 2, 2)
    for i, v in enumerate(Bt.views(S)):
        assert torch.equal(v, ref[i]), i


@pytest.mark.gpu
def test_nanodisk_decisions_match_oracle():
    from sck_amd.sampler import Sampler
    s = nanodisk(4.5, 3)
    B = 8
    A = Sampler(s, B, DEV, seed=7)
    A.set_temperature(torch.linspace(0.3, 2.0, B))
    S = A.init_spins(seed=5)
    os_ = O.OracleSystem(s)
    beta = A.beta.cpu().numpy()
    for c in range(len(A.classes)):
        x0 = S.cpu().numpy()
        A.state[1] = 0
        A.sweep(S, 1, classes=(c, c + 1))
        ref, _, mg = O.chain(os_, x0, beta, 7, 0, 1, class_range=(c, c + 1), margins=True)
        got = S.cpu().numpy()
        bad = np.abs(got - ref).max(axis=1) > 1e-5
        near = np.abs(mg.reshape(bad.shape)) < 1e-5
        assert not np.any(bad & ~near), f"class {c}: {int((bad & ~near).sum())} decisions differ"
        S.copy_(torch.as_tensor(ref, device=DEV))


@pytest.mark.gpu
def test_single_spin_site_anisotropy_distribution_is_exact():
    """One Heisenberg spin with a site anisotropy Ks along a tilted axis a and a field h along z:
    <s_z> and <(s.a)^2> under e^{beta (h s_z + Ks (s.a)^2)} by quadrature over the sphere."""
    from sck_amd.sampler import Sampler
    s, ax, Ks, h, T = _site_anisotropy_spin()
    B = 4096
    A = Sampler(s, B, DEV, seed=4)
    A.set_temperature(T)
    S = A.init_spins(seed=1)
    A.sweep(S, 200)
    sz, sa = [], []
    axt = torch.as_tensor(ax, device=DEV)
    for _ in range(40):
        A.sweep(S, 5)
        v = S[:, :, 0, 0, 0, 0].double()
        sz.append(v[:, 2].mean().item())
        sa.append(((v @ axt) ** 2).mean().item())
    ez, ea = _sphere_average(ax, Ks, h, T)
    for series, exact in ((sz, ez), (sa, ea)):
        assert abs(np.mean(series) - exact) < 3 * np.std(series) / math.sqrt(len(series)) + 2e-3, exact


@pytest.mark.gpu
def test_full_size_nanodisk_batch_matches_oracle():
    """The C4 batch at its real size (64 nanodisks, radius U[8, 40], thickness U{4..12} from default_rng(2024),
    16 replicas each, T in [0.3, 1.8]): after each class launch of sweep 0, four (system, replica) pairs, the
    largest disk among them, match the oracle's decisions from the same state."""
    from sck_amd.hamiltonian import nanodisk_set
    systems = nanodisk_set(64, 8.0, 40.0, 4, 12, seed=2024)
    Bt = BatchSampler(systems, 16, DEV, seed=0, schedule=Schedule())
    temps = np.linspace(0.3, 1.8, 16).astype(np.float32)
    Bt.set_temperature([temps] * len(systems))
    S = Bt.init_spins(seed=0)
    big = int(np.argmax([s.n_active for s in systems]))
    picks = [(0, 0), (17, 5), (63, 15), (big, 9)]
    beta = Bt.beta.cpu().numpy()
    for c in range(len(Bt.classes)):
        views = Bt.views(S)
        x0 = {p: views[p[0]][[p[1]]].cpu().numpy() for p in picks}
        Bt.state[1] = 0
        Bt.sweep(S, 1, classes=(c, c + 1))
        views = Bt.views(S)
        for i, r in picks:
            off = int(Bt.rep_off[i])
            ref, _, mg = O.chain(O.OracleSystem(systems[i]), x0[(i, r)], beta[off:off + 16], Bt.seeds[i], 0, 1,
                                 reps=[r], class_range=(c, c + 1), margins=True)
            got = views[i][[r]].cpu().numpy()
            bad = np.abs(got - ref).max(axis=1) > 1e-5
            near = np.abs(mg.reshape(bad.shape)) < 1e-5
            assert not np.any(bad & ~near), f"class {c} system {i} replica {r}"
            views[i][r] = torch.as_tensor(ref[0], device=DEV)
