"""The tiled sweep kernel (sck_tiled): one launch per sweep, each workgroup running every colour class of one
(replica, tile) in LDS with a halo of K bond reaches.

CPU: the plan (tile, halo per axis) and the layout policy.
GPU: whole sweeps through the tiled kernel equal the per-class launches (which the oracle pins class by class)
spin for spin, on 2-D systems that cover periodic and open axes, masks, fields, uniaxial anisotropy, extra
(DMI) bond matrices, several sublattices, partial tiles and the cone proposal, and on the full XY triangular
2048^2 batch (B = 2048).
"""

import numpy as np
import pytest
import torch

from sck_amd import crystals as cr
from sck_amd import hipgen as G
from sck_amd import system as sy
from sck_amd.sampler import choose_layout

DEV = "cuda"


def _dmi_square(L1, L2, D=0.3):
    """Heisenberg square lattice, open along x, periodic along y, interfacial DMI on the first shell, a field
    and a disc-shaped hole in the mask."""
    c = cr.make("square")
    extra = {}
    for i, bd in enumerate(c.bonds):
        r = np.array(list(bd.delta) + [0], float)
        E = np.zeros((3, 3))
        # D (r x z) . (S_i x S_j) written as a 3x3 bond matrix
        dvec = D * np.cross(r, [0.0, 0.0, 1.0])
        E[1, 2], E[2, 1] = dvec[0], -dvec[0]
        E[2, 0], E[0, 2] = dvec[1], -dvec[1]
        extra[i] = E
    yy, xx = np.mgrid[0:L2, 0:L1]
    mask = ((xx - L1 / 2) ** 2 + (yy - L2 / 2) ** 2 > 36)[None, None]
    return sy.System("Heis square DMI open-x masked", 3, c, (L1, L2), (False, True), {1: 1.0}, extra,
                     h=(0.0, 0.1, 0.2), mask=mask)


def _xy_triangular_open(L1, L2):
    """XY triangular lattice, open along x, periodic along y, with a field and a disc-shaped hole in the mask."""
    yy, xx = np.mgrid[0:L2, 0:L1]
    mask = ((xx - L1 / 2) ** 2 + (yy - L2 / 2) ** 2 > 30)[None, None]
    return sy.System("XY triangular open-x masked", 2, cr.make("triangular"), (L1, L2), (False, True), {1: 1.0},
                     h=(0.1, 0.2), mask=mask)


SYSTEMS = {
    "xy_triangular_98": (lambda: sy.xy("triangular", 98), {}, None),
    "xy_triangular_96_3col": (lambda: sy.xy("triangular", 96), {}, None),
    "heis_square_100x90_partial": (
        lambda: sy.System("Heis square", 3, cr.make("square"), (100, 90), (True, True), {1: 1.0},
                          h=(0.0, 0.0, 0.3), uniaxial=(0.2, (0.0, 0.0, 1.0))), {}, None),
    "heis_kagome_afm_small_tiles": (lambda: sy.heisenberg("kagome", 60, J=-1.0), {}, (24, 24)),
    "heis_dmi_open_masked": (lambda: _dmi_square(90, 80), {}, None),
    "xy_square_cone": (lambda: sy.xy("square", 80), {"proposal": "cone"}, None),
    "heis_triangular_barker": (lambda: sy.heisenberg("triangular", 96, J=-1.0), {"barker": True}, (32, 48)),
    # four colours, float4 rows with a cell's two components side by side in LDS, partial tiles
    "xy_triangular_100": (lambda: sy.xy("triangular", 100), {}, None),
    "xy_triangular_open_masked_cone": (lambda: _xy_triangular_open(100, 64), {"proposal": "cone"}, None),
    # 512-thread workgroups: taller tiles (up to 80 KB of LDS), partial tiles along both axes
    "xy_triangular_300_512_threads": (lambda: sy.xy("triangular", 300), {}, (128, 64, 512)),
    "heis_square_100x90_512_threads": (
        lambda: sy.System("Heis square", 3, cr.make("square"), (100, 90), (True, True), {1: 1.0},
                          h=(0.0, 0.0, 0.3), uniaxial=(0.2, (0.0, 0.0, 1.0))), {}, (64, 48, 512)),
}


def test_tiled_plan_geometry():
    s = sy.xy("triangular", 2048)
    g = G._Gen(s, G.MODE_F32, 2048, "uniform")
    assert len(g.classes) == 4                     # 2048 mod 3 != 0: the seam condition forces four colours
    assert g.tiled_plan() == (128, 32, 4, 4, 1, 1)
    assert G._Gen(s, G.MODE_F32, 2048, "uniform", tile_dims=(64, 64)).tiled_plan() == (64, 64, 4, 4, 1, 1)
    # 512 threads: a 128 x 64 tile and its halo (136 x 72 cells, 78 KB) fit the larger LDS budget
    g512 = G._Gen(s, G.MODE_F32, 2048, "uniform", tile_dims=(128, 64, 512))
    assert g512.tiled_plan() == (128, 64, 4, 4, 1, 1) and g512.TILE_THREADS == 512
    assert not G.fits_tiled(s, G.MODE_F32, (128, 64))                    # 78 KB exceeds the default 64 KB
    m = G.generate(s, G.MODE_F32, 2048, tiled=True)
    assert m.tile_dims == (128, 32, 512) and "amdgpu_waves_per_eu(1, 6)" in m.src
    assert G.generate(s, G.MODE_F32, 2048, tiled=True, tile_dims=(128, 32, 256)).tiled.block[0] == 256
    m512 = G.generate(s, G.MODE_F32, 2048, tiled=True, tile_dims=(128, 64, 512))
    assert m512.tile_dims == (128, 64, 512) and m512.tiled.block[0] == 512
    assert "amdgpu_waves_per_eu(1, 4)" in m512.src                        # two workgroups of 8 waves per CU
    with pytest.raises(ValueError, match="256, 512 or 1024"):
        G._Gen(s, G.MODE_F32, 2048, "uniform", tile_dims=(128, 64, 384))
    # Heisenberg triangular, three colours: the tile is rounded to the colouring period and to float4 columns
    assert G._Gen(sy.heisenberg("triangular", 300), G.MODE_F32, 8, "uniform").tiled_plan()[:2] == (120, 30)
    assert not G.fits_tiled(sy.xy("triangular", 12), G.MODE_F32)          # smaller than a tile and its halo
    assert not G.fits_tiled(sy.heisenberg("bcc", 32), G.MODE_F32)         # 3-D
    assert not G.fits_tiled(sy.ising("square", 256), G.MODE_INT8)         # Ising
    k = sy.heisenberg("kagome", 60, J=-1.0)
    assert not G.fits_tiled(k, G.MODE_F32, (64, 64)) and G.fits_tiled(k, G.MODE_F32, (24, 24))   # LDS budget
    assert choose_layout(s, G.MODE_F32, 2048) == ("reference", True)
    assert choose_layout(sy.xy("square", 256), G.MODE_F32, 64) == ("replica", False)   # cache-resident
    assert choose_layout(sy.xy("square", 256), G.MODE_F32, 16) == ("reference", True)
    assert choose_layout(sy.xy("square", 256), G.MODE_F32, 64, layout="replica") == ("replica", False)
    with pytest.raises(ValueError, match="tiled sweep"):
        choose_layout(sy.heisenberg("bcc", 16), G.MODE_F32, 64, tiled=True)
    src = G.generate(s, G.MODE_F32, 2048, tiled=True).src
    assert "sck_tiled" in src and "__shared__" in src


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(SYSTEMS))
def test_tiled_sweeps_equal_class_launches(name):
    from sck_amd.sampler import Sampler
    make, kw, dims = SYSTEMS[name]
    s = make()
    if kw.pop("barker", False):
        s.acceptance = "barker"
    B = 24
    print(f"[gpu] tiled == class launches {name} {s.box} B={B}", flush=True)
    if kw.get("proposal") == "cone":
        kw["width"] = np.linspace(0.3, 3.0, B).astype(np.float32)
    T = Sampler(s, B, DEV, seed=5, tiled=True, tile_dims=dims, resident=False, **kw)
    R = Sampler(s, B, DEV, seed=5, tiled=False, layout="reference", resident=False, **kw)
    assert T.use_tiled and T.module.tiled is not None and not R.use_tiled and R.module.tiled is None
    temps = torch.linspace(0.3, 2.5, B)
    T.set_temperature(temps)
    R.set_temperature(temps)
    S1 = T.init_spins(seed=3)
    S2 = S1.clone()
    for k in range(3):
        T.sweep(S1, 1)
        R.sweep(S2, 1)
        assert torch.equal(S1, S2), f"sweep {k}: {int((S1 != S2).sum())} components differ"
    assert int(T.state[1]) == 3
    T.sweep(S1, 2)                                 # several sweeps in one call
    R.sweep(S2, 2)
    assert torch.equal(S1, S2)
    for k in (2, 3):                               # captured in graphs: an even and an odd number of sweeps
        rp = T.graph(S1, k)
        rp()
        rp()
        R.sweep(S2, 2 * k)
        torch.cuda.synchronize()
        assert torch.equal(S1, S2), k
    # through the C-ABI without a workspace (args.dst NULL) the same module runs its class launches
    import ctypes as C

    from sck_amd import _lib as L
    a = T._args(S1)
    assert not a.dst
    L.check(L.lib().sck_sweep(T._handle, C.byref(a), 1, 0, -1, C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    R.sweep(S2, 1)
    torch.cuda.synchronize()
    assert torch.equal(S1, S2) and int(T.state[1]) == int(R.state[1])
    if s.mask is not None:
        m = torch.as_tensor(s.site_mask(), device=DEV)
        assert torch.all(S1.norm(dim=1)[:, ~m] == 0)


@pytest.mark.gpu
def test_tiled_full_size_xy_equals_class_launches():
    """XY triangular 2048^2, B = 2048 (16 temperatures x 128 in [0.7, 1.1]): one tiled sweep equals the four class
    launches spin for spin over the whole batch."""
    import bench
    from sck_amd.sampler import Sampler
    s, _, B, temps, _ = bench._make("xy")
    print(f"[gpu] full-size tiled xy: B={B}", flush=True)
    T = Sampler(s, B, DEV, seed=0)
    assert T.use_tiled and T.layout == "reference"
    R = Sampler(s, B, DEV, seed=0, tiled=False, layout="reference")
    T.set_temperature(temps)
    R.set_temperature(temps)
    S1 = bench.device_spins(T, seed=0)
    S2 = S1.clone()
    X0 = S1[:64].clone()
    T.sweep(S1, 1)
    R.sweep(S2, 1)
    torch.cuda.synchronize()
    for r0 in range(0, B, 128):
        assert torch.equal(S1[r0:r0 + 128], S2[r0:r0 + 128]), f"replicas {r0}..{r0 + 127} differ"
    changed = (S1[:64] != X0).any(dim=1).float().mean().item()
    assert 0.05 < changed < 0.95


@pytest.mark.gpu
def test_graph_pair_equals_graph_replays():
    """Sampler.graph_pair: two graphs with the spins and the workspace exchanged (no copy back after the odd sweep
    count), replayed in turn, run the chain that graph() replays run, and observe the same values after every
    replay; after an odd number of replays the state is in the workspace."""
    from sck_amd.sampler import Sampler
    make, kw, dims = SYSTEMS["xy_triangular_98"]
    s = make()
    B = 16
    print(f"[gpu] graph pair {s.name} {s.box}", flush=True)
    A = Sampler(s, B, DEV, seed=5, tiled=True, tile_dims=dims, resident=False, **kw)
    R = Sampler(s, B, DEV, seed=5, tiled=True, tile_dims=dims, resident=False, **kw)
    for X in (A, R):
        X.set_temperature(torch.linspace(0.4, 1.6, B))
    S1 = A.init_spins(seed=3)
    S2 = S1.clone()
    there, back, W = A.graph_pair(S1, 1, observe=True)
    assert W.data_ptr() != S1.data_ptr()
    ref = R.graph(S2, 1, observe=True)
    for i in range(5):
        (back if i % 2 else there)()
        ref()
        torch.cuda.synchronize()
        cur = S1 if i % 2 else W
        assert torch.equal(cur, S2), i
        Ea, Ma = A.observed()
        Er, Mr = R.observed()
        assert torch.equal(Ea, Er) and torch.equal(Ma, Mr), i
    assert int(A.state[1]) == int(R.state[1]) == 5
