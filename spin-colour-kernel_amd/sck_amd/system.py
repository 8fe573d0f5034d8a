"""A system: spin model, crystal, finite box and couplings.

Mirrors SCK/system.py:19-204. Spins use the component-planar, basis-planar layout
S[B, n, b, L3, L2, L1] with x fastest (SCK/system.py:3-6); masked sites hold the zero vector, are never
updated, and therefore contribute nothing to a neighbour's local field.

Hamiltonian (reduced units, kB = 1):

    H = - sum_<ij> J_s S_i.S_j - sum_<ij> S_i^T A_ij S_j - sum_i h.S_i
        + sum_i K1_i (sx^2 sy^2 + sx^2 sz^2 + sy^2 sz^2) + K2 sum_i (sx sy sz)^2      (reference terms)
        - Du sum_i (S_i.e)^2 - sum_i Ku_i (S_i.n_i)^2                                    (terms added here)

The two uniaxial terms are not in the reference. `uniaxial = (Du, e)` is one easy axis e for every site
(Du > 0 favours S parallel to e); `site_ku` / `site_axis` give every site its own strength and unit axis,
which is how a surface anisotropy normal to a nanodisk's surface is expressed (sck_amd.hamiltonian).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import crystals as cr
from .colouring import Colouring, colour_grid, search, seam_ok, verify_on_box

__all__ = ["System", "ising", "xy", "heisenberg", "nanodot", "matrix", "TC"]


@dataclass
class System:
    name: str
    n: int                                   # spin components: 1 Ising, 2 XY, 3 Heisenberg
    crystal: cr.Crystal
    box: tuple[int, ...]                     # (L1, L2[, L3]) cells
    periodic: tuple[bool, ...]
    J: dict[int, float]                      # isotropic exchange per shell
    extra: dict[int, np.ndarray] = field(default_factory=dict)   # bond index -> extra n x n matrix (DMI ...)
    h: tuple[float, ...] | None = None       # uniform field
    cubic_k1: np.ndarray | None = None       # per-site cubic anisotropy K1, shape (b, L3, L2, L1)
    cubic_k2: float = 0.0                    # (sx sy sz)^2 coefficient
    mask: np.ndarray | None = None           # bool (b, L3, L2, L1); False = no spin
    colouring: Colouring | None = None
    T_ref: float | None = None               # reference critical temperature, when one exists
    acceptance: str = "metropolis"           # "metropolis": u < exp(-beta dE); "barker": heat-bath rule
    uniaxial: tuple[float, tuple[float, ...]] | None = None     # (Du, unit axis e): - Du (S.e)^2 per site
    site_ku: np.ndarray | None = None        # (b, L3, L2, L1) per-site uniaxial strength Ku_i
    site_axis: np.ndarray | None = None      # (b, L3, L2, L1, n) per-site unit axis n_i

    def __post_init__(self):
        if self.n not in (1, 2, 3):
            raise ValueError("n must be 1, 2 or 3")
        self.box = tuple(int(v) for v in self.box)
        self.periodic = tuple(bool(v) for v in self.periodic)
        if len(self.box) != self.crystal.D or len(self.periodic) != self.crystal.D:
            raise ValueError("box and periodic must have one entry per lattice axis")
        if self.cubic_k1 is not None and self.n != 3:
            raise ValueError("cubic anisotropy needs n = 3")
        if (self.uniaxial is not None or self.site_ku is not None) and self.n == 1:
            raise ValueError("uniaxial anisotropy needs n = 2 or 3")
        if (self.site_ku is None) != (self.site_axis is None):
            raise ValueError("site_ku and site_axis go together")
        if self.uniaxial is not None:
            Du, e = self.uniaxial
            e = np.asarray(e, float).reshape(-1)
            if e.size != self.n or not np.isclose(np.linalg.norm(e), 1.0):
                raise ValueError(f"the uniaxial axis must be a unit vector of {self.n} components")
            self.uniaxial = (float(Du), tuple(float(v) for v in e))
        if self.site_ku is not None:
            self.site_ku = np.asarray(self.site_ku, np.float32).reshape(self.shape)
            self.site_axis = np.asarray(self.site_axis, np.float32).reshape(self.shape + (self.n,))
        if self.acceptance not in ("metropolis", "barker"):
            raise ValueError("acceptance must be 'metropolis' or 'barker'")
        if self.colouring is None:
            self.colouring = search(self.crystal, self.box, self.periodic)
            if self.colouring is None:
                raise ValueError(f"{self.name}: no seam-compatible linear colouring for box {self.box}")
        if not seam_ok(self.colouring, self.box, self.periodic):
            raise ValueError(f"{self.name}: colouring violates the seam condition (8) on box {self.box}")
        if not verify_on_box(self.colouring, self.crystal, self.box, self.periodic):
            raise ValueError(f"{self.name}: colouring is not admissible on the box")

    # ---------------------------------------------------------------- geometry
    @property
    def L1(self) -> int:
        return int(self.box[0])

    @property
    def L2(self) -> int:
        return int(self.box[1])

    @property
    def L3(self) -> int:
        return int(self.box[2]) if self.crystal.D == 3 else 1

    @property
    def b(self) -> int:
        return self.crystal.b

    @property
    def shape(self) -> tuple[int, int, int, int]:
        return (self.b, self.L3, self.L2, self.L1)

    @property
    def nsite(self) -> int:
        return self.b * self.L3 * self.L2 * self.L1

    def site_mask(self) -> np.ndarray:
        return np.ones(self.shape, bool) if self.mask is None else np.asarray(self.mask, bool).reshape(self.shape)

    @property
    def n_active(self) -> int:
        return int(self.site_mask().sum())

    def bond_matrix(self, i: int) -> np.ndarray:
        bd = self.crystal.bonds[i]
        M = self.J.get(bd.shell, 0.0) * np.eye(self.n)
        if i in self.extra:
            M = M + self.extra[i]
        return M

    def colour_map(self) -> np.ndarray:
        """Colour of every site, shape (b, L3, L2, L1)."""
        return colour_grid(self.colouring, self.b, self.box)

    def classes(self) -> list[tuple[int, int]]:
        """(colour, sublattice) classes in sweep order: colour-major, then sublattice, empty ones skipped
        (SCK/sampler.py:127-129)."""
        cm, m = self.colour_map(), self.site_mask()
        return [(c, a) for c in range(self.colouring.n_colours) for a in range(self.b)
                if np.any((cm[a] == c) & m[a])]


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


# -------------------------------------------------------------------- matrix
def ising(crystal: str, L: int, J: float = 1.0, shells=(1,), T_ref=None) -> System:
    c = cr.make(crystal, tuple(shells))
    return System(f"Ising {crystal}", 1, c, (L,) * c.D, (True,) * c.D, {s: J for s in shells}, T_ref=T_ref)


def xy(crystal: str, L: int, J: float = 1.0, T_ref=None) -> System:
    c = cr.make(crystal)
    return System(f"XY {crystal}", 2, c, (L,) * c.D, (True,) * c.D, {1: J}, T_ref=T_ref)


def heisenberg(crystal: str, L: int, J: float = 1.0, T_ref=None, label=None) -> System:
    c = cr.make(crystal)
    return System(label or f"Heisenberg {crystal}", 3, c, (L,) * c.D, (True,) * c.D, {1: J}, T_ref=T_ref)


def _sc_neighbour_count(mask3: np.ndarray) -> np.ndarray:
    """Active first neighbours of every site of a (L, N, N) sc mask: periodic in x and y, open in z."""
    L = mask3.shape[0]
    m = mask3.astype(np.int64)
    k = np.roll(m, -1, 2) + np.roll(m, 1, 2) + np.roll(m, -1, 1) + np.roll(m, 1, 1)
    up = np.zeros_like(m)
    dn = np.zeros_like(m)
    if L > 1:
        up[:-1] = m[1:]
        dn[1:] = m[:-1]
    return k + up + dn


def nanodot(Rd: float = 18.30, L: int = 5, Jex=1.0, Kan1=0.177, Kan2=0.034, KanS=0.139, Hex=0.483,
            KDM=0.88, gamma_deg=90.0, shells=(1,), J_shells=None) -> System:
    """The nanodot system of the reference (SCK/system.py:117-171): an sc Heisenberg disc of radius Rd and
    L layers in an open box, bulk DMI on the in-plane first-shell bonds, cubic anisotropy with KanS on
    surface sites (fewer than six in-disc first neighbours), and a Zeeman field Hex (cos g, 0, sin g)."""
    shells = tuple(sorted(set(shells) | {1}))
    c = cr.make("sc", shells)
    J = {1: Jex, **{int(k): float(v) for k, v in (J_shells or {}).items() if int(k) in shells and int(k) != 1}}
    L2 = int(Rd) + 1
    N = 2 * L2 + 1
    xs = np.arange(-L2, L2 + 1)
    xg, yg = np.meshgrid(xs, xs, indexing="xy")
    disk = (xg ** 2 + yg ** 2) <= Rd ** 2               # [y, x]
    mask = np.broadcast_to(disk, (L, N, N)).copy()[None]  # (1, L, N, N)
    surface = mask & (_sc_neighbour_count(mask[0])[None] < 6)
    k1 = np.where(surface, KanS, Kan1).astype(np.float32) * mask
    extra = {}
    for i, bd in enumerate(c.bonds):
        E = np.zeros((3, 3))
        if bd.delta == (1, 0, 0):
            E[1, 2], E[2, 1] = KDM, -KDM
        elif bd.delta == (-1, 0, 0):
            E[1, 2], E[2, 1] = -KDM, KDM
        elif bd.delta == (0, 1, 0):
            E[0, 2], E[2, 0] = -KDM, KDM
        elif bd.delta == (0, -1, 0):
            E[0, 2], E[2, 0] = KDM, -KDM
        if E.any():
            extra[i] = E
    g = np.deg2rad(gamma_deg)
    return System("Heisenberg nanodot", 3, c, (N, N, L), (False, False, False), J, extra,
                  h=(Hex * float(np.cos(g)), 0.0, Hex * float(np.sin(g))), cubic_k1=k1, cubic_k2=Kan2,
                  mask=mask)


#: reference critical temperatures (J = kB = 1), SCK/system.py:174-184
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


def matrix(small: bool = True) -> list[System]:
    """The 13 systems of the reference's exactness matrix at a small size per crystal."""
    return [
        ising("square", 8, T_ref=TC["Ising square"]),
        ising("triangular", 6, T_ref=TC["Ising triangular"]),
        ising("sc", 4, T_ref=TC["Ising sc"]),
        ising("bcc", 4, T_ref=TC["Ising bcc"]),
        ising("fcc", 4, T_ref=TC["Ising fcc"]),
        xy("square", 8, T_ref=0.893),
        xy("sc", 4, T_ref=TC["XY sc"]),
        heisenberg("sc", 4, T_ref=TC["Heisenberg sc"]),
        heisenberg("bcc", 4, T_ref=TC["Heisenberg bcc"]),
        heisenberg("fcc", 4, T_ref=None),
        heisenberg("triangular", 6, J=-1.0, T_ref=0.285, label="Heisenberg triangular AFM"),
        heisenberg("kagome", 6, J=-1.0, T_ref=None, label="Heisenberg kagome AFM"),
        nanodot(Rd=3.0, L=3),
    ]
