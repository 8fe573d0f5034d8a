"""Spin systems: a model (Ising, XY, Heisenberg), a crystal with a basis, a finite box and the couplings.

The field names and their order are the reference's System (SCK/system.py:19-34), so code written against it runs
unchanged; the spin layout is S[B, n, b, L3, L2, L1] with x fastest (SCK/system.py:3-6). A masked site holds the
zero vector and is never updated, so it adds nothing to a neighbour's local field.

Energy (reduced units, kB = 1):

    H = - sum_<ij> J_s S_i.S_j - sum_<ij> S_i^T A_ij S_j - sum_i h.S_i
        + sum_i K1_i (sx^2 sy^2 + sx^2 sz^2 + sy^2 sz^2) + K2 sum_i (sx sy sz)^2      (reference terms)
        - Du sum_i (S_i.e)^2 - sum_i Ku_i (S_i.n_i)^2                                    (added here)

The two uniaxial terms have no counterpart in the reference: `uniaxial = (Du, e)` is one easy axis for every site
(Du > 0 favours S parallel to e), and `site_ku` / `site_axis` give each site its own strength and axis, which is
how sck_amd.hamiltonian expresses an anisotropy normal to a nanodisk's surface.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import crystals as cr
from .colouring import Colouring, colour_grid, search, seam_ok, verify_on_box

__all__ = ["System", "as_system", "ising", "xy", "heisenberg", "nanodot", "matrix", "TC"]

_MODELS_N = (1, 2, 3)
_RULES = ("metropolis", "barker")


def _check_terms(s: "System") -> None:
    """The model-dependent terms: which n each one needs, and the shapes of the per-site arrays."""
    if s.n not in _MODELS_N:
        raise ValueError("n must be 1, 2 or 3")
    if len(s.box) != s.crystal.D or len(s.periodic) != s.crystal.D:
        raise ValueError("box and periodic must have one entry per lattice axis")
    if s.cubic_k1 is not None and s.n != 3:
        raise ValueError("cubic anisotropy needs n = 3")
    if s.n == 1 and (s.uniaxial is not None or s.site_ku is not None):
        raise ValueError("uniaxial anisotropy needs n = 2 or 3")
    if (s.site_ku is None) is not (s.site_axis is None):
        raise ValueError("site_ku and site_axis go together")
    if s.acceptance not in _RULES:
        raise ValueError("acceptance must be 'metropolis' or 'barker'")


def _normalise_terms(s: "System") -> None:
    if s.uniaxial is not None:
        strength, axis = s.uniaxial
        e = np.asarray(axis, float).ravel()
        if e.size != s.n or abs(float(e @ e) - 1.0) > 1e-6:
            raise ValueError(f"the uniaxial axis must be a unit vector of {s.n} components")
        s.uniaxial = (float(strength), tuple(map(float, e)))
    if s.site_ku is not None:
        s.site_ku = np.asarray(s.site_ku, np.float32).reshape(s.shape)
        s.site_axis = np.asarray(s.site_axis, np.float32).reshape(s.shape + (s.n,))


def _attach_colouring(s: "System") -> None:
    """Search a colouring when none is given, then require the seam condition (8) and admissibility on the box."""
    if s.colouring is None:
        s.colouring = search(s.crystal, s.box, s.periodic)
        if s.colouring is None:
            raise ValueError(f"{s.name}: no seam-compatible linear colouring for box {s.box}")
    if not seam_ok(s.colouring, s.box, s.periodic):
        raise ValueError(f"{s.name}: colouring violates the seam condition (8) on box {s.box}")
    if not verify_on_box(s.colouring, s.crystal, s.box, s.periodic):
        raise ValueError(f"{s.name}: colouring is not admissible on the box")


@dataclass
class System:
    name: str
    n: int                                   # spin components: 1 Ising, 2 XY, 3 Heisenberg
    crystal: cr.Crystal
    box: tuple[int, ...]                     # cells along x, y (and z)
    periodic: tuple[bool, ...]
    J: dict[int, float]                      # isotropic exchange of each neighbour shell
    extra: dict[int, np.ndarray] = field(default_factory=dict)   # bond index -> n x n matrix added to J (DMI)
    h: tuple[float, ...] | None = None       # uniform field, n components
    cubic_k1: np.ndarray | None = None       # per-site cubic K1, (b, L3, L2, L1)
    cubic_k2: float = 0.0                    # coefficient of (sx sy sz)^2
    mask: np.ndarray | None = None           # (b, L3, L2, L1) bool, False where no spin sits
    colouring: Colouring | None = None
    T_ref: float | None = None               # a reference (critical) temperature, where one is known
    acceptance: str = "metropolis"           # "metropolis" or "barker" (heat-bath rule)
    uniaxial: tuple[float, tuple[float, ...]] | None = None     # (Du, unit axis e): - Du (S.e)^2 on every site
    site_ku: np.ndarray | None = None        # (b, L3, L2, L1) per-site uniaxial strength Ku_i
    site_axis: np.ndarray | None = None      # (b, L3, L2, L1, n) per-site unit axis n_i

    def __post_init__(self):
        self.box = tuple(map(int, self.box))
        self.periodic = tuple(map(bool, self.periodic))
        _check_terms(self)
        _normalise_terms(self)
        _attach_colouring(self)

    # --------------------------------------------------------------------------------- geometry
    @property
    def b(self) -> int:
        return self.crystal.b

    @property
    def L1(self) -> int:
        return self.box[0]

    @property
    def L2(self) -> int:
        return self.box[1]

    @property
    def L3(self) -> int:
        return self.box[2] if self.crystal.D == 3 else 1

    @property
    def shape(self) -> tuple[int, int, int, int]:
        """(b, L3, L2, L1): the site axes of the spin layout."""
        return (self.b, self.L3, self.L2, self.L1)

    @property
    def nsite(self) -> int:
        return int(np.prod(self.shape))

    def site_mask(self) -> np.ndarray:
        if self.mask is None:
            return np.ones(self.shape, bool)
        return np.asarray(self.mask, bool).reshape(self.shape)

    @property
    def n_active(self) -> int:
        return int(np.count_nonzero(self.site_mask()))

    # --------------------------------------------------------------------------------- couplings
    def bond_matrix(self, i: int) -> np.ndarray:
        """The n x n coupling of directed bond i: J of its shell times the identity, plus its extra matrix."""
        shell = self.crystal.bonds[i].shell
        M = np.eye(self.n) * self.J.get(shell, 0.0)
        return M + self.extra[i] if i in self.extra else M

    def colour_map(self) -> np.ndarray:
        """Colour of every site, (b, L3, L2, L1)."""
        return colour_grid(self.colouring, self.b, self.box)

    def classes(self) -> list[tuple[int, int]]:
        """The update classes in sweep order: (colour, sublattice) pairs holding an active site, colour-major
        (SCK/sampler.py:127-129)."""
        present = (self.colour_map()[..., None] == np.arange(self.colouring.n_colours)) & self.site_mask()[..., None]
        held = present.reshape(self.b, -1, self.colouring.n_colours).any(axis=1)      # (b, colours)
        return [(c, a) for c in range(self.colouring.n_colours) for a in range(self.b) if held[a, c]]


def as_system(obj) -> System:
    """`obj` as a System. An object that only carries the reference System's fields (SCK/system.py:19-34: name,
    n, crystal, box, periodic, J, extra, h, cubic_k1, cubic_k2, mask, colouring, T_ref, acceptance), such as
    the reference's own System, gets the terms it lacks as None (no uniaxial or per-site anisotropy); its
    crystal and colouring are used as they are (their bond table and (m, K, o) fix the kernel), and the class
    list is derived from its colours and mask as SCK/sampler.py:127-129 does."""
    if isinstance(obj, System):
        return obj
    for f in ("name", "n", "crystal", "box", "periodic", "J", "colouring"):
        if not hasattr(obj, f):
            raise ValueError(f"not a spin system: no field {f!r}")
    return System(str(obj.name), int(obj.n), obj.crystal, tuple(obj.box), tuple(obj.periodic), dict(obj.J),
                  dict(getattr(obj, "extra", None) or {}), getattr(obj, "h", None), getattr(obj, "cubic_k1", None),
                  float(getattr(obj, "cubic_k2", 0.0) or 0.0), getattr(obj, "mask", None), obj.colouring,
                  getattr(obj, "T_ref", None), getattr(obj, "acceptance", "metropolis"),
                  getattr(obj, "uniaxial", None), getattr(obj, "site_ku", None), getattr(obj, "site_axis", None))


# ------------------------------------------------------------------------------------- standard systems
def _uniform(kind: str, n: int, crystal: str, L: int, J: float, shells=(1,), T_ref=None, label=None) -> System:
    """A periodic L^D box of one crystal with the same J on every requested shell."""
    c = cr.make(crystal, tuple(shells))
    return System(label or f"{kind} {crystal}", n, c, (L,) * c.D, (True,) * c.D, dict.fromkeys(shells, J),
                  T_ref=T_ref)


def ising(crystal: str, L: int, J: float = 1.0, shells=(1,), T_ref=None) -> System:
    return _uniform("Ising", 1, crystal, L, J, shells, T_ref)


def xy(crystal: str, L: int, J: float = 1.0, T_ref=None) -> System:
    return _uniform("XY", 2, crystal, L, J, T_ref=T_ref)


def heisenberg(crystal: str, L: int, J: float = 1.0, T_ref=None, label=None) -> System:
    return _uniform("Heisenberg", 3, crystal, L, J, T_ref=T_ref, label=label)


def nanodot(Rd: float = 18.30, L: int = 5, Jex=1.0, Kan1=0.177, Kan2=0.034, KanS=0.139, Hex=0.483,
            KDM=0.88, gamma_deg=90.0, shells=(1,), J_shells=None) -> System:
    """The reference's nanodot (SCK/system.py:117-171): a Heisenberg disc of radius Rd and L layers on the sc
    crystal inside an open box of (2 (int(Rd) + 1) + 1)^2 x L cells, with

      * exchange Jex on the first shell and J_shells {shell: J} on any further `shells`;
      * bulk DMI of strength KDM on the in-plane first-shell bonds (A_lm = KDM sum_k r_k eps_klm for the unit
        bond vector r; the out-of-plane bonds carry none);
      * cubic anisotropy Kan1, KanS on surface sites (fewer than six in-disc first neighbours), and Kan2;
      * a Zeeman field Hex (cos g, 0, sin g).

    The disc stays at least one cell clear of the box edge, so an open box holds the same Hamiltonian as a
    periodic one."""
    from .hamiltonian import dmi, surface
    shells = tuple(sorted(set(shells) | {1}))
    c = cr.make("sc", shells)
    J = {1: Jex}
    J.update({int(k): float(v) for k, v in (J_shells or {}).items() if int(k) in shells and int(k) != 1})
    half = int(Rd) + 1
    N = 2 * half + 1
    d = np.arange(N) - half
    disc = d[None, :] ** 2 + d[:, None] ** 2 <= Rd ** 2                      # [y, x]
    mask = np.repeat(disc[None, None], L, axis=1)                            # (1, L, N, N)
    box, open_ = (N, N, L), (False, False, False)
    k1 = (np.where(surface(c, box, open_, mask), KanS, Kan1) * mask).astype(np.float32)
    g = np.deg2rad(gamma_deg)
    return System("Heisenberg nanodot", 3, c, box, open_, J, dmi(c, KDM),
                  h=(Hex * float(np.cos(g)), 0.0, Hex * float(np.sin(g))), cubic_k1=k1, cubic_k2=Kan2,
                  mask=mask)


#: reference critical temperatures (J = kB = 1) of the reference's systems, SCK/system.py:174-184
TC = {
    "Ising square": 2.0 / np.log(1 + np.sqrt(2)),
    "Ising triangular": 4.0 / np.log(3.0),
    "Ising sc": 4.511523,
    "Ising bcc": 6.355,
    "Ising fcc": 9.797132,
    "XY sc": 2.20184,
    "Heisenberg sc": 1.4430,
    "Heisenberg bcc": 2.0542,
}

# the reference's exactness matrix (SCK/system.py:187-204): (model, crystal, L, J, T_ref, label)
_MATRIX = (
    ("ising", "square", 8, 1.0, TC["Ising square"], None),
    ("ising", "triangular", 6, 1.0, TC["Ising triangular"], None),
    ("ising", "sc", 4, 1.0, TC["Ising sc"], None),
    ("ising", "bcc", 4, 1.0, TC["Ising bcc"], None),
    ("ising", "fcc", 4, 1.0, TC["Ising fcc"], None),
    ("xy", "square", 8, 1.0, 0.893, None),
    ("xy", "sc", 4, 1.0, TC["XY sc"], None),
    ("heisenberg", "sc", 4, 1.0, TC["Heisenberg sc"], None),
    ("heisenberg", "bcc", 4, 1.0, TC["Heisenberg bcc"], None),
    ("heisenberg", "fcc", 4, 1.0, None, None),
    ("heisenberg", "triangular", 6, -1.0, 0.285, "Heisenberg triangular AFM"),
    ("heisenberg", "kagome", 6, -1.0, None, "Heisenberg kagome AFM"),
)


def matrix(small: bool = True) -> list[System]:
    """The 13 systems of the reference's exactness matrix at a small size per crystal, the nanodot last."""
    make = {"ising": lambda c, L, J, T, lab: ising(c, L, J, T_ref=T),
            "xy": lambda c, L, J, T, lab: xy(c, L, J, T_ref=T),
            "heisenberg": lambda c, L, J, T, lab: heisenberg(c, L, J, T_ref=T, label=lab)}
    return [make[m](c, L, J, T, lab) for m, c, L, J, T, lab in _MATRIX] + [nanodot(Rd=3.0, L=3)]
