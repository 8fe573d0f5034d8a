"""Hamiltonian builders: state parameters -> System (SCK/simulate.py:109-240), plus the nanodisk systems with
surface-normal anisotropy used by the batched sweeps.

Restated from the reference: bulk-type (Bloch) DMI A_lm = D sum_k r_k eps_klm on in-plane first-shell bonds
(SCK/simulate.py:153-168), the surface as active sites with fewer active first-shell neighbours than a bulk site
of their basis index (SCK/simulate.py:171-195), and `build_system` for the state keys J, D, K1, Ks, K2, H,
H_theta, H_phi with the open-box fallback (SCK/simulate.py:198-240).

New here (no counterpart in the reference): a uniaxial anisotropy along each surface site's outward normal,
E = - Ks sum_{i in surface} (S_i . n_i)^2 with n_i the normalised sum of the bond vectors to its missing
first-shell neighbours, carried by System.site_ku / site_axis.
"""

from __future__ import annotations

import numpy as np

from . import crystals as cr
from .system import System

__all__ = ["dmi", "surface", "surface_normals", "build_system", "nanodisk", "nanodisk_set", "MODELS", "STATE_KEYS"]

MODELS = {"ising": 1, "xy": 2, "heisenberg": 3}


def _first_shell(crystal: cr.Crystal):
    """(bond index, bond, Cartesian bond vector padded to 3 components) of every first-shell bond."""
    for i, bd in enumerate(crystal.bonds):
        if bd.shell == 1:
            offset = np.asarray(bd.delta, float) + crystal.basis[bd.alpha2] - crystal.basis[bd.alpha]
            r = np.zeros(3)
            r[:crystal.D] = offset @ crystal.vectors
            yield i, bd, r


def dmi(crystal: cr.Crystal, D: float) -> dict[int, np.ndarray]:
    """Bulk-type (Bloch) DMI on the in-plane first-shell bonds. With u the unit bond vector the bond matrix is
    A_lm = D sum_k u_k eps_klm, the skew matrix for which S_i . A S_j = D u . (S_i x S_j)."""
    out = {}
    for i, _, r in _first_shell(crystal):
        if abs(r[2]) > 1e-9:                       # out-of-plane bonds carry none
            continue
        x, y, z = r / np.linalg.norm(r)
        out[i] = D * np.array([[0.0, z, -y], [-z, 0.0, x], [y, -x, 0.0]])
    return out


def _partner(a: np.ndarray, delta, periodic, D: int) -> np.ndarray:
    """Array p of a's shape (L3, L2, L1) with p[z, y, x] = a at (x, y, z) + delta: the index wraps across a
    periodic face, and beyond an open face the value is 0."""
    out = a
    for k in range(D):
        d = int(delta[k])
        if d == 0:
            continue
        axis = 2 - k                                # lattice axis k is array axis 2 - k of (z, y, x)
        L = a.shape[axis]
        idx = np.arange(L) + d
        if periodic[k]:
            out = np.take(out, idx % L, axis=axis)
        else:
            inside = ((idx >= 0) & (idx < L)).reshape([L if j == axis else 1 for j in range(3)])
            out = np.take(out, np.clip(idx, 0, L - 1), axis=axis) * inside
    return out


def surface(crystal: cr.Crystal, box, periodic, mask: np.ndarray) -> np.ndarray:
    """Active sites with fewer active first-shell neighbours than a bulk site of their basis index has bonds."""
    bonds = [bd for _, bd, _ in _first_shell(crystal)]
    bulk = np.bincount([bd.alpha for bd in bonds], minlength=crystal.b).reshape(-1, 1, 1, 1)
    active = np.zeros(mask.shape, int)
    for bd in bonds:
        active[bd.alpha] += _partner(mask[bd.alpha2].astype(int), bd.delta, periodic, crystal.D)
    return mask & (active < bulk)


def surface_normals(crystal: cr.Crystal, box, periodic, mask: np.ndarray) -> np.ndarray:
    """Outward unit normal of every surface site: the normalised sum of the bond vectors to its missing
    first-shell neighbours (the first missing bond when they cancel), zero elsewhere. Shape mask.shape + (3,)."""
    acc = np.zeros(mask.shape + (3,))
    first = np.zeros(mask.shape + (3,))
    has_first = np.zeros(mask.shape, bool)
    for _, bd, r in _first_shell(crystal):
        present = _partner(mask[bd.alpha2].astype(int), bd.delta, periodic, crystal.D).astype(bool)
        missing = mask[bd.alpha] & ~present
        acc[bd.alpha] += missing[..., None] * r
        new = missing & ~has_first[bd.alpha]
        first[bd.alpha][new] = r
        has_first[bd.alpha] |= missing
    norm = np.linalg.norm(acc, axis=-1, keepdims=True)
    n = np.where(norm > 1e-9, acc / np.maximum(norm, 1e-30), first / np.maximum(np.linalg.norm(first, axis=-1, keepdims=True), 1e-30))
    return n * surface(crystal, box, periodic, mask)[..., None]


STATE_KEYS = ("name", "J", "D", "K1", "Ks", "K2", "H", "H_theta", "H_phi")


def _number(theta: dict, key: str, default: float = 0.0) -> float:
    """theta[key] as a float; an absent key or None reads as `default`."""
    v = theta.get(key)
    return float(default) if v is None else float(v)


def _field(n: int, H: float, theta_deg: float, phi_deg: float):
    """The field of strength H along polar angle theta and azimuth phi (degrees), or None for H = 0. Ising takes
    H itself and XY the in-plane direction phi."""
    if not H:
        return None
    t, p = np.deg2rad(theta_deg), np.deg2rad(phi_deg)
    if n == 1:
        comps = (H,)
    elif n == 2:
        comps = (H * np.cos(p), H * np.sin(p))
    else:
        comps = (H * np.sin(t) * np.cos(p), H * np.sin(t) * np.sin(p), H * np.cos(t))
    return tuple(float(c) for c in comps)


def _site_k1(c: cr.Crystal, cells, box, periodic, mask, K1: float, Ks: float) -> np.ndarray:
    """Per-site K1 (float32): Ks on the surface sites of the active region, K1 on its other sites, 0 outside."""
    active = np.ones(cells, bool) if mask is None else mask
    return (np.where(surface(c, box, periodic, active), Ks, K1) * active).astype(np.float32)


def build_system(model: str, crystal: str, box, periodic, theta: dict, mask=None, log=None) -> System:
    """The System of one state theta (the parameters of SCK/simulate.py:198-240: J shells, bulk DMI D, K1 with Ks
    on surface sites, K2, field H at angles H_theta, H_phi). A periodic box with no admissible colouring is
    built open instead, with a message to `log`."""
    key = str(model).strip().lower()
    if key not in MODELS:
        raise ValueError(f"unknown model {model!r}; known: {', '.join(MODELS)}")
    n = MODELS[key]
    J = {int(s): float(v) for s, v in (theta.get("J") or {1: 1.0}).items()}
    D, K1, K2, H = (_number(theta, k) for k in ("D", "K1", "K2", "H"))
    Ks = _number(theta, "Ks", K1)                  # the surface takes the bulk value unless Ks is given
    if n != 3 and any(_number(theta, k) for k in ("D", "K1", "Ks", "K2")):
        raise ValueError(f"D, K1, Ks and K2 need the Heisenberg model; {key} carries J and H only")
    unknown = set(theta) - set(STATE_KEYS)
    if unknown:
        raise ValueError(f"unknown parameters {sorted(unknown)}; known: {', '.join(STATE_KEYS[1:])}")
    c = cr.make(str(crystal), tuple(sorted({1} | {s for s, v in J.items() if v != 0.0})))
    box = tuple(int(v) for v in box)
    cells = (c.b, box[2] if c.D == 3 else 1, box[1], box[0])
    m = None if mask is None else np.asarray(mask, bool)
    if m is not None and m.shape != cells:
        raise ValueError(f"mask shape {m.shape} does not match the box {box} of {c.name}")
    terms = dict(extra=dmi(c, D) if (n == 3 and D) else {}, h=_field(n, H, _number(theta, "H_theta"),
                 _number(theta, "H_phi")), cubic_k2=K2, mask=m)
    name = str(theta.get("name") or f"{key} {c.name}")
    anisotropic = n == 3 and bool(K1 or Ks or K2)
    periodic = tuple(bool(v) for v in periodic)
    tries = [periodic] + ([(False,) * c.D] if any(periodic) else [])
    for i, per in enumerate(tries):
        k1 = _site_k1(c, cells, box, per, m, K1, Ks) if anisotropic else None
        try:
            return System(name, n, c, box, per, J, cubic_k1=k1, **terms)
        except ValueError as e:
            if i == len(tries) - 1 or "colouring" not in str(e):
                raise
            (log or print)(f"  no admissible colouring for the periodic box {box}: using open boundaries")


def nanodisk(radius: float, thickness: int, J1: float = 1.0, Ks: float = 0.3, D: float = 0.1,
             name: str | None = None) -> System:
    """Heisenberg nanodisk on the sc crystal: a cylinder of `radius` sites and `thickness` layers in an open box
    (2 ceil(radius) + 1)^2 x thickness, exchange J1, bulk DMI D on the in-plane first-shell bonds and a uniaxial
    anisotropy Ks along the outward surface normal of every surface site."""
    c = cr.make("sc")
    L = int(np.ceil(radius))
    N = 2 * L + 1
    xs = np.arange(-L, L + 1)
    xg, yg = np.meshgrid(xs, xs, indexing="xy")
    disk = (xg ** 2 + yg ** 2) <= radius ** 2
    mask = np.broadcast_to(disk, (thickness, N, N)).copy()[None]
    box, per = (N, N, int(thickness)), (False, False, False)
    normals = surface_normals(c, box, per, mask)
    ku = np.where(np.linalg.norm(normals, axis=-1) > 0, Ks, 0.0).astype(np.float32)
    return System(name or f"nanodisk R={radius:g} t={thickness}", 3, c, box, per, {1: J1}, dmi(c, D),
                  mask=mask, site_ku=ku, site_axis=normals.astype(np.float32))


def nanodisk_set(count: int = 64, rmin: float = 8.0, rmax: float = 40.0, tmin: int = 4, tmax: int = 12,
                 seed: int = 2024, **kw) -> list[System]:
    """`count` nanodisks with radius uniform in [rmin, rmax] and thickness uniform in {tmin .. tmax}, drawn from
    numpy's default_rng(seed)."""
    rng = np.random.default_rng(seed)
    radii = rng.uniform(rmin, rmax, size=count)
    thick = rng.integers(tmin, tmax + 1, size=count)
    return [nanodisk(float(r), int(t), **kw) for r, t in zip(radii, thick)]
