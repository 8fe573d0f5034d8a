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

__all__ = ["dmi", "surface", "surface_normals", "build_system", "nanodisk", "nanodisk_set", "MODELS"]

MODELS = {"ising": 1, "xy": 2, "heisenberg": 3}


def dmi(crystal: cr.Crystal, D: float) -> dict[int, np.ndarray]:
    """Bulk-type DMI matrices A_lm = D sum_k r_k eps_klm on the in-plane first-shell bonds (r the unit bond)."""
    eps = np.zeros((3, 3, 3))
    for i, j, k in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        eps[i, j, k], eps[i, k, j] = 1.0, -1.0
    out = {}
    for i, bd in enumerate(crystal.bonds):
        if bd.shell != 1:
            continue
        r = np.zeros(3)
        r[:crystal.D] = (np.asarray(bd.delta, float) + crystal.basis[bd.alpha2] - crystal.basis[bd.alpha]) @ crystal.vectors
        if abs(r[2]) > 1e-9:
            continue
        r /= np.linalg.norm(r)
        out[i] = D * np.einsum("k,klm->lm", r, eps)
    return out


def _shift(src: np.ndarray, delta, box, periodic, D):
    """src (L3, L2, L1) moved so that element [z, y, x] holds src at (x, y, z) + delta, 0 outside an open box."""
    out = src
    valid = np.ones(src.shape, bool)
    axes = {0: 2, 1: 1, 2: 0}
    for k in range(D):
        d = delta[k]
        if d == 0:
            continue
        out = np.roll(out, -d, axis=axes[k])
        if not periodic[k]:
            idx = np.arange(box[k]) + d
            ok = (idx >= 0) & (idx < box[k])
            shp = [1, 1, 1]
            shp[axes[k]] = box[k]
            valid = valid & ok.reshape(shp)
    return np.where(valid, out, 0)


def surface(crystal: cr.Crystal, box, periodic, mask: np.ndarray) -> np.ndarray:
    """Active sites with fewer active first-shell neighbours than a bulk site of their basis index."""
    full = {a: 0 for a in range(crystal.b)}
    count = np.zeros(mask.shape, int)
    for bd in crystal.bonds:
        if bd.shell != 1:
            continue
        full[bd.alpha] += 1
        count[bd.alpha] += _shift(mask[bd.alpha2].astype(int), bd.delta, box, periodic, crystal.D)
    bulk = np.stack([np.full(mask.shape[1:], full[a]) for a in range(crystal.b)])
    return mask & (count < bulk)


def surface_normals(crystal: cr.Crystal, box, periodic, mask: np.ndarray) -> np.ndarray:
    """Outward unit normal of every surface site: the normalised sum of the bond vectors to its missing
    first-shell neighbours (the first missing bond when they cancel), zero elsewhere. Shape mask.shape + (3,)."""
    acc = np.zeros(mask.shape + (3,))
    first = np.zeros(mask.shape + (3,))
    has_first = np.zeros(mask.shape, bool)
    for bd in crystal.bonds:
        if bd.shell != 1:
            continue
        r = np.zeros(3)
        r[:crystal.D] = (np.asarray(bd.delta, float) + crystal.basis[bd.alpha2] - crystal.basis[bd.alpha]) @ crystal.vectors
        present = _shift(mask[bd.alpha2].astype(int), bd.delta, box, periodic, crystal.D).astype(bool)
        missing = mask[bd.alpha] & ~present
        acc[bd.alpha] += missing[..., None] * r
        new = missing & ~has_first[bd.alpha]
        first[bd.alpha][new] = r
        has_first[bd.alpha] |= missing
    norm = np.linalg.norm(acc, axis=-1, keepdims=True)
    n = np.where(norm > 1e-9, acc / np.maximum(norm, 1e-30), first / np.maximum(np.linalg.norm(first, axis=-1, keepdims=True), 1e-30))
    return n * surface(crystal, box, periodic, mask)[..., None]


def _f(theta, key, default=0.0) -> float:
    v = theta.get(key, default)
    return float(default if v is None else v)


def build_system(model: str, crystal: str, box, periodic, theta: dict, mask=None, log=None) -> System:
    """The System of one state theta; a periodic box without an admissible colouring falls back to open."""
    model = str(model).strip().lower()
    if model not in MODELS:
        raise ValueError(f"unknown model {model!r}; known: {', '.join(MODELS)}")
    n = MODELS[model]
    J = {int(k): float(v) for k, v in (theta.get("J") or {1: 1.0}).items()}
    D, K1, K2, H = _f(theta, "D"), _f(theta, "K1"), _f(theta, "K2"), _f(theta, "H")
    Ks = _f(theta, "Ks", K1)
    if n != 3 and any(_f(theta, k) for k in ("D", "K1", "Ks", "K2")):
        raise ValueError(f"D, K1, Ks and K2 need the Heisenberg model; {model} carries J and H only")
    unknown = set(theta) - {"name", "J", "D", "K1", "Ks", "K2", "H", "H_theta", "H_phi"}
    if unknown:
        raise ValueError(f"unknown parameters {sorted(unknown)}; known: J, D, K1, Ks, K2, H, H_theta, H_phi")
    shells = tuple(sorted({s for s, v in J.items() if v != 0.0} | {1}))
    c = cr.make(str(crystal), shells)
    box = tuple(int(v) for v in box)
    periodic = tuple(bool(v) for v in periodic)
    m = None if mask is None else np.asarray(mask, bool)
    if m is not None and m.shape != (c.b, box[2] if c.D == 3 else 1, box[1], box[0]):
        raise ValueError(f"mask shape {m.shape} does not match the box {box} of {c.name}")
    t, p = np.deg2rad(_f(theta, "H_theta")), np.deg2rad(_f(theta, "H_phi"))
    h = None
    if H:
        h = {1: (H,), 2: (H * np.cos(p), H * np.sin(p)),
             3: (H * np.sin(t) * np.cos(p), H * np.sin(t) * np.sin(p), H * np.cos(t))}[n]
        h = tuple(float(v) for v in h)
    kw = dict(extra=dmi(c, D) if (n == 3 and D) else {}, h=h, cubic_k2=K2, mask=m)
    name = theta.get("name") or f"{model} {c.name}"

    def make(per):
        k1 = None
        if n == 3 and (K1 or Ks or K2):
            mm = np.ones((c.b, box[2] if c.D == 3 else 1, box[1], box[0]), bool) if m is None else m
            k1 = (np.where(surface(c, box, per, mm), Ks, K1) * mm).astype(np.float32)
        return System(str(name), n, c, box, per, J, cubic_k1=k1, **kw)
    try:
        return make(periodic)
    except ValueError as e:
        if not any(periodic) or "colouring" not in str(e):
            raise
        (log or print)(f"  no admissible colouring for the periodic box {box}: using open boundaries")
        return make((False,) * c.D)


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
