"""The reference-side binding (sck_amd/mi355x.py, the file a maintainer adds as SCK/mi355x.py) driven by objects
that carry exactly the reference's System fields (SCK/system.py:19-34) and Sampler attributes
(SCK/sampler.py:40-131) -- no classes(), uniaxial, site_ku or Ising table of their own.

CPU: the generator accepts such a System and emits the same source as for this build's System; hipRTC compiles
it (no device needed); the binding refuses a non-HIP device.
GPU: sweeps through the binding reproduce the reference's own trajectories (the ising_square8, xy_square4 and
bitwise_indep_square8 fixtures), with the Ising table rewritten in place on a temperature change.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import pytest
import torch

from sck_amd import _lib as L
from sck_amd import crystals as cr
from sck_amd import hipgen as G
from sck_amd import system as sy

GOLDEN = Path(__file__).resolve().parent / "golden"
WORD = 32


@dataclass
class RefSystem:
    """The reference System's fields and methods only (SCK/system.py:19-98)."""

    name: str
    n: int
    crystal: object
    box: tuple
    periodic: tuple
    J: dict
    extra: dict = field(default_factory=dict)
    h: tuple | None = None
    cubic_k1: np.ndarray | None = None
    cubic_k2: float = 0.0
    mask: np.ndarray | None = None
    colouring: object = None
    T_ref: float | None = None
    acceptance: str = "metropolis"

    @property
    def L1(self):
        return int(self.box[0])

    @property
    def L2(self):
        return int(self.box[1])

    @property
    def L3(self):
        return int(self.box[2]) if self.crystal.D == 3 else 1

    @property
    def b(self):
        return self.crystal.b

    @property
    def shape(self):
        return (self.b, self.L3, self.L2, self.L1)

    def site_mask(self):
        return np.ones(self.shape, bool) if self.mask is None else self.mask.astype(bool)

    @property
    def n_active(self):
        return int(self.site_mask().sum())

    def colour_map(self):
        out = np.zeros(self.shape, np.int64)
        for a in range(self.b):
            for z in range(self.L3):
                for y in range(self.L2):
                    for x in range(self.L1):
                        out[a, z, y, x] = self.colouring.colour((x, y, z)[: self.crystal.D], a)
        return out


def ref_system(s: sy.System) -> RefSystem:
    """The reference-shaped twin of a System (its crystal and colouring objects carry the reference's fields)."""
    names = [f.name for f in dataclasses.fields(RefSystem)]
    return RefSystem(**{k: getattr(s, k) for k in names})


class RefSampler:
    """The attributes the reference Sampler holds after __init__ (SCK/sampler.py:40-131) and its
    set_temperature (SCK/sampler.py:134-146), plus the one line that calls the binding."""

    def __init__(self, system, batch, device="cuda", seed=0, packed=None, bitwise=False, independent=False,
                 spin_dtype=None):
        from sck_amd.mi355x import MI355XKernel
        from sck_amd.sampler import p_fixed, p_table
        self.sys = system
        self.bitwise, self.independent = bool(bitwise), bool(independent)
        if self.bitwise:
            packed = True
        self.packed = (system.n == 1 and batch >= WORD) if packed is None else bool(packed)
        self.spin_dtype = spin_dtype or (torch.int8 if system.n == 1 else torch.float32)
        self.B = int(batch)
        self.device = torch.device(device)
        self.W = -(-self.B // WORD) if self.packed else self.B
        self.Bp = self.W * WORD if self.packed else self.B
        m = system.site_mask()
        self.mask = torch.as_tensor(m.astype(np.int8), device=self.device)
        k1 = system.cubic_k1 if system.cubic_k1 is not None else np.zeros(m.shape, np.float32)
        self.k1 = torch.as_tensor(np.asarray(k1, np.float32), device=self.device)
        self.state = torch.tensor([int(seed), 0], dtype=torch.int64, device=self.device)
        self.beta = torch.ones(self.Bp, dtype=torch.float32, device=self.device)
        if self.bitwise:
            self._table = p_fixed if self.independent else p_table
            self.ptab = torch.as_tensor(self._table(system, np.ones(self.W)), device=self.device)
        self._mi355x = MI355XKernel(self)

    def set_temperature(self, T):
        T = torch.as_tensor(T, dtype=torch.float32, device=self.device).reshape(-1)
        beta = (1.0 / T).expand(self.B).contiguous() if T.numel() == 1 else (1.0 / T).contiguous()
        if self.Bp > self.B:
            beta = torch.cat([beta, beta[-1:].expand(self.Bp - self.B)])
        if self.bitwise:
            bwd = beta.reshape(self.W, WORD)
            self.ptab.copy_(torch.as_tensor(self._table(self.sys, bwd[:, 0].double().cpu().numpy()),
                                            device=self.device))
        self.beta.copy_(beta)
        self._mi355x.set_temperature(self)

    def sweep(self, S, n=1):
        return self._mi355x.sweep(self, S, n)


# ------------------------------------------------------------------------------------------------ CPU
SYSTEMS = [lambda: sy.ising("square", 8), lambda: sy.xy("square", 4), lambda: sy.nanodot(Rd=2.0, L=2),
           lambda: sy.heisenberg("kagome", 3, J=-1.0)]


@pytest.mark.parametrize("i", range(len(SYSTEMS)))
def test_generator_accepts_the_reference_system(i):
    s = SYSTEMS[i]()
    r = ref_system(s)
    assert not hasattr(r, "classes") and not hasattr(r, "uniaxial") and not hasattr(r, "site_ku")
    a = sy.as_system(r)
    # the class list of SCK/sampler.py:127-129, derived from the reference-shaped object's own colour map
    m, cm = r.site_mask(), r.colour_map()
    want = [(c, al) for c in range(r.colouring.n_colours) for al in range(r.b) if np.any((cm[al] == c) & m[al])]
    assert a.classes() == want == s.classes()
    assert a.uniaxial is None and a.site_ku is None and a.site_axis is None
    mode = G.MODE_INT8 if s.n == 1 else G.MODE_F32
    g1, g2 = G.generate(r, mode, 4), G.generate(s, mode, 4)
    assert g1.src == g2.src and g1.kernels == g2.kernels and g1.classes == want
    L.precompile(g1.src)                       # hipRTC for gfx950, no device needed


def test_generator_refuses_an_object_that_is_not_a_system():
    with pytest.raises(ValueError, match="not a spin system"):
        G.generate(object(), G.MODE_F32, 4)


def test_binding_refuses_a_cpu_device():
    from sck_amd.mi355x import MI355XKernel

    class Cpu:
        sys = ref_system(sy.ising("square", 4))
        device = "cpu"
    with pytest.raises(ValueError, match="HIP device"):
        MI355XKernel(Cpu())


# ------------------------------------------------------------------------------------------------ GPU
def _recipe_system(name):
    from recipes import RECIPES
    rc = RECIPES[name]
    s = eval(rc["system"], {"sy": sy, "cr": cr, "np": np})
    if rc.get("acceptance"):
        s.acceptance = rc["acceptance"]
    return rc, ref_system(s)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ising_square8", "xy_square4", "bitwise_indep_square8"])
def test_binding_reproduces_reference_trajectories(name):
    print(f"[gpu] reference binding {name}", flush=True)
    rc, r = _recipe_system(name)
    z = np.load(GOLDEN / f"{name}.npz")
    kw = rc.get("sampler", {})
    A = RefSampler(r, rc["B"], "cuda", seed=rc["seed"], packed=kw.get("packed"), bitwise=kw.get("bitwise", False),
                   independent=kw.get("independent", False))
    # first at another temperature, then the recipe's: the binding's table must follow set_temperature
    A.set_temperature(torch.full((rc["B"],), 0.37))
    A.set_temperature(torch.as_tensor(rc["T"], dtype=torch.float32))
    assert np.array_equal(A.beta.cpu().numpy(), z["beta"])
    S = torch.as_tensor(z["S0"]).to("cuda").contiguous()
    for k in range(rc["sweeps"]):
        A.sweep(S, 1)
        got = S.cpu().numpy()
        if r.n == 1:
            assert np.array_equal(got, z["S"][k]), f"sweep {k}: {int((got != z['S'][k]).sum())} spins differ"
        else:
            assert np.allclose(got, z["S"][k], rtol=0, atol=1e-5), f"sweep {k}"
    assert int(A.state[1]) == rc["sweeps"]
    if not A.packed:
        E = A._mi355x.energy(A, S).cpu().numpy()
        M = A._mi355x.magnetization(A, S).cpu().numpy()
        assert np.allclose(E, z["E"], rtol=1e-9, atol=1e-4) and np.allclose(M, z["M"], rtol=1e-9, atol=1e-4)


@pytest.mark.gpu
def test_binding_float64_storage_follows_the_reference_trajectory():
    """A reference Sampler holding float64 spins (spin_dtype=torch.float64; its kernel loads them as float32 and
    stores the cast, SCK/generator.py:214, 310-319): the binding runs float64 storage and follows the
    reference's xy_square4 trajectory."""
    from sck_amd import hipgen
    rc, r = _recipe_system("xy_square4")
    z = np.load(GOLDEN / "xy_square4.npz")
    A = RefSampler(r, rc["B"], "cuda", seed=rc["seed"], spin_dtype=torch.float64)
    assert A._mi355x.mode == hipgen.MODE_F64
    A.set_temperature(torch.as_tensor(rc["T"], dtype=torch.float32))
    S = torch.as_tensor(z["S0"]).to("cuda", torch.float64).contiguous()
    for k in range(rc["sweeps"]):
        A.sweep(S, 1)
        assert S.dtype == torch.float64
        assert np.allclose(S.cpu().numpy(), z["S"][k], rtol=0, atol=1e-5), f"sweep {k}"
    E = A._mi355x.energy(A, S).cpu().numpy()
    assert np.allclose(E, z["E"], rtol=1e-9, atol=1e-4)


@pytest.mark.gpu
def test_binding_reference_chain_and_graph_equal_its_sweeps():
    """Through the binding: the sequential chain equals the coloured sweep, and a captured graph of k sweeps
    replayed twice equals 2k eager sweeps, on a reference-shaped Heisenberg kagome system."""
    r = ref_system(sy.heisenberg("kagome", 3, J=-1.0))
    A = RefSampler(r, 3, "cuda", seed=4)
    A.set_temperature(torch.tensor([0.4, 1.0, 2.0]))
    from sck_amd.sampler import Sampler
    S = Sampler(sy.as_system(r), 3, "cuda", seed=4).init_spins(seed=1)
    S1, S2, S3 = S.clone(), S.clone(), S.clone()
    A.sweep(S1, 2)
    A.state[1] = 0
    A._mi355x.reference_sweep(A, S2, 2)
    assert torch.equal(S1, S2)
    A.state[1] = 0
    rp = A._mi355x.graph(A, S3, 1)
    rp()
    rp()
    torch.cuda.synchronize()
    assert torch.equal(S3, S1) and int(A.state[1]) == 2
