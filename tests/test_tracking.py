"""Running totals of E and M kept by the tiled sweep kernel (Sampler.track): every (replica, tile) workgroup sums
the energy and spin changes of its accepted moves in float64 (a wave reduction, then one partial per workgroup),
and a fold kernel adds one sweep's partials to the per-replica totals in a fixed order.

Bar: right after track() (an anchor) the totals equal observe()'s exactly; after further sweeps they agree with
observe() to 1e-6 relative (the energy) and 1e-7 of the site count (the magnetization). The drift is the fp32
rounding of each move's local field and spin change, a random walk over the accepted moves (measured: about
5e-9 relative after one sweep of a 98^2 lattice). Sweeps run eagerly and in graph replays, on periodic, open,
masked, Barker, cone-proposal, several-sublattice and partial-tile systems, and on the full XY triangular
2048^2 batch. A sampler without the tiled sweep refuses to track.
"""

import numpy as np
import pytest
import torch

from sck_amd import system as sy

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _close(A, S):
    E, M = A.observe(S)
    Et, Mt = A.tracked()
    N = float(A.sys.n_active)
    rel = (Et - E).abs().max().item() / E.abs().max().item()
    dm = (Mt - M).abs().max().item() / N
    return rel, dm


def _systems():
    import test_tiled as TT
    return {k: v for k, v in TT.SYSTEMS.items()}


@pytest.mark.parametrize("name", ["xy_triangular_98", "heis_square_100x90_partial", "heis_kagome_afm_small_tiles",
                                  "heis_dmi_open_masked", "xy_square_cone", "heis_triangular_barker"])
def test_running_totals_follow_the_observables(name):
    from sck_amd.sampler import Sampler
    make, kw, dims = _systems()[name]
    kw = dict(kw)
    s = make()
    if kw.pop("barker", False):
        s.acceptance = "barker"
    B = 16
    print(f"[gpu] running totals {name} {s.box}", flush=True)
    if kw.get("proposal") == "cone":
        kw["width"] = np.linspace(0.3, 3.0, B).astype(np.float32)
    A = Sampler(s, B, DEV, seed=2, tiled=True, tile_dims=dims, resident=False, **kw)
    A.set_temperature(torch.linspace(0.4, 2.5, B))
    S = A.init_spins(seed=1)
    E0, M0 = A.observe(S)
    Et, Mt = A.track(S)
    assert torch.equal(Et, E0) and torch.equal(Mt, M0)
    A.sweep(S, 1)
    rel, dm = _close(A, S)
    assert rel < 1e-6 and dm < 1e-7, (rel, dm)
    A.sweep(S, 7)                                  # an odd count: the copy back after the last tiled sweep
    rp = A.graph(S, 4)
    for _ in range(3):
        rp()
    torch.cuda.synchronize()
    rel, dm = _close(A, S)
    assert rel < 1e-6 and dm < 1e-7, (rel, dm)
    E1, _ = A.observe(S)
    assert not torch.equal(E1, E0)                 # the totals moved with the chain
    held = A._track.clone()
    A.untrack()
    A.sweep(S, 2)                                  # no longer tracked: the totals stay where they were
    torch.cuda.synchronize()
    assert torch.equal(A._track, held)
    with pytest.raises(ValueError, match="track"):
        A.tracked()


def test_tracking_needs_the_tiled_sweep():
    from sck_amd.sampler import Sampler
    A = Sampler(sy.heisenberg("bcc", 4), 64, DEV, seed=1, layout="replica")
    S = A.init_spins(seed=1)
    with pytest.raises(ValueError, match="tiled sweep"):
        A.track(S)


def test_fused_series_equals_observed_series():
    """observe.series with fused running totals (re-anchored every 7 measurements) gives the series the
    observation graphs give, within the running totals' bar."""
    from sck_amd.observe import series
    from sck_amd.sampler import Sampler
    s = sy.xy("triangular", 98)
    B = 32
    A = Sampler(s, B, DEV, seed=4, tiled=True, resident=False)
    R = Sampler(s, B, DEV, seed=4, tiled=True, resident=False)
    for X in (A, R):
        X.set_temperature(torch.linspace(0.6, 1.6, B))
    S1 = A.init_spins(seed=3)
    S2 = S1.clone()
    E1, M1, _ = series(A, S1, 30, every=2, fused=True, anchor_every=7)
    E2, M2, _ = series(R, S2, 30, every=2)
    assert torch.equal(S1, S2)
    assert np.allclose(E1, E2, rtol=1e-6, atol=0) and np.allclose(M1, M2, rtol=0, atol=1e-7)


def test_full_size_xy_running_totals():
    """XY triangular 2048^2, B = 2048: after 6 tiled sweeps the running totals agree with observe()."""
    import bench
    from sck_amd.sampler import Sampler
    s, _, B, T, _ = bench._make("xy")
    print(f"[gpu] full-size running totals xy B={B}", flush=True)
    A = Sampler(s, B, DEV, seed=0)
    assert A.use_tiled
    A.set_temperature(T)
    S = bench.device_spins(A, seed=0)
    A.track(S)
    A.sweep(S, 6)
    rel, dm = _close(A, S)
    assert rel < 1e-6 and dm < 1e-7, (rel, dm)
