"""Linear colourings gamma(c, alpha) = K c + o_alpha (mod m) and their checks.

Mirrors SCK/colouring.py:21-161. A colouring maps a site into the finite abelian group
Z_m1 x ... x Z_mr (invariant-factor form m1 | m2 | ...); its colour index is the mixed-radix number of that
group element (SCK/colouring.py:30-36). Admissibility (no bond joins two sites of one colour,
SCK/colouring.py:62-71) makes every (colour, sublattice) class a set of mutually unbonded sites, so one class
can be updated in parallel and equals a sequential sweep over it. The seam condition
(SCK/colouring.py:74-79) keeps the colouring consistent across a periodic box edge.

`search` enumerates candidates in the same order as the reference (group order n_c ascending, groups in
the order `abelian_groups` yields them, K columns then offsets in itertools.product order), because the
colouring it returns fixes the class order and therefore the Markov chain. The admissibility test itself is
vectorised over all bonds; `verify_on_box` and `colour_map` are vectorised over all sites, which the
reference does with Python loops (SCK/colouring.py:103-120, SCK/system.py:89-98).
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass

import numpy as np

from .crystals import Crystal

__all__ = ["Colouring", "abelian_groups", "admissible", "seam_ok", "search", "verify_on_box",
           "colour_grid", "clique_number"]


@dataclass(frozen=True)
class Colouring:
    m: tuple[int, ...]                 # group orders, invariant-factor form
    K: tuple[tuple[int, ...], ...]     # r rows, D columns
    o: tuple[tuple[int, ...], ...]     # b rows, r columns

    @property
    def n_colours(self) -> int:
        return int(np.prod(self.m)) if self.m else 1

    def colour(self, c, alpha) -> int:
        """Mixed-radix index of gamma(c, alpha) in [0, n_colours)."""
        idx = 0
        for r, mr in enumerate(self.m):
            g = (sum(self.K[r][k] * int(c[k]) for k in range(len(c))) + self.o[alpha][r]) % mr
            idx = idx * mr + g
        return idx

    def describe(self) -> str:
        grp = " x ".join(f"Z{m}" for m in self.m) or "trivial"
        return f"{grp}; K={[list(r) for r in self.K]}; o={[list(x) for x in self.o]}"


def abelian_groups(n: int) -> list[tuple[int, ...]]:
    """Every abelian group of order n as invariant factors (m1 | m2 | ...), in the reference's order:
    the first factor ascending, each later factor a multiple of the one before."""
    if n == 1:
        return [()]
    out: list[tuple[int, ...]] = []

    # depth-first, smallest divisor first: the enumeration order of SCK/colouring.py:41-57
    def rec(rem: int, prev: int, acc: tuple[int, ...]) -> None:
        if rem == 1:
            out.append(acc)
            return
        for d in range(2, rem + 1):
            if rem % d == 0 and d % prev == 0:
                rec(rem // d, d, acc + (d,))

    rec(n, 1, ())
    return [g for g in out if all(g[i + 1] % g[i] == 0 for i in range(len(g) - 1))]


def _bond_arrays(crystal: Crystal):
    alpha = np.array([bd.alpha for bd in crystal.bonds], np.int64)
    alpha2 = np.array([bd.alpha2 for bd in crystal.bonds], np.int64)
    delta = np.array([bd.delta for bd in crystal.bonds], np.int64).reshape(-1, crystal.D)
    return alpha, alpha2, delta


def admissible(col: Colouring, crystal: Crystal) -> bool:
    """K delta + o_alpha2 - o_alpha != 0 (mod m) for every bond."""
    if not crystal.bonds:
        return True
    alpha, alpha2, delta = _bond_arrays(crystal)
    zero = np.ones(len(alpha), bool)
    for r, mr in enumerate(col.m):
        K = np.asarray(col.K[r], np.int64)
        o = np.asarray([x[r] for x in col.o], np.int64)
        zero &= ((delta @ K + o[alpha2] - o[alpha]) % mr) == 0
    return not bool(zero.any())


def seam_ok(col: Colouring, box: tuple[int, ...], periodic: tuple[bool, ...]) -> bool:
    """Every periodic axis k has K[:, k] L_k = 0 (mod m)."""
    for k, (Lk, per) in enumerate(zip(box, periodic)):
        if per and any((col.K[r][k] * Lk) % mr for r, mr in enumerate(col.m)):
            return False
    return True


def search(crystal: Crystal, box: tuple[int, ...] | None = None, periodic: tuple[bool, ...] | None = None,
           max_colours: int = 32) -> Colouring | None:
    """The first admissible linear colouring in increasing n_c, seam-compatible with `box` when given."""
    D, b = crystal.D, crystal.b
    per = tuple(periodic) if periodic is not None else (True,) * D
    alpha, alpha2, delta = _bond_arrays(crystal) if crystal.bonds else (None, None, None)
    for n in range(1, max_colours + 1):
        for m in abelian_groups(n):
            r = len(m)
            elems = list(itertools.product(*[range(mi) for mi in m])) if r else [()]
            mvec = np.asarray(m, np.int64)
            for cols in itertools.product(elems, repeat=D):
                K = tuple(tuple(cols[k][i] for k in range(D)) for i in range(r))
                probe = Colouring(m, K, ((0,) * r,) * b)
                if box is not None and not seam_ok(probe, tuple(box), per):
                    continue
                Kd = (delta @ np.asarray(K, np.int64).T) if (r and alpha is not None) else None   # (nbond, r)
                for offs in itertools.product(elems, repeat=b - 1):
                    o = (tuple([0] * r),) + tuple(offs)
                    if alpha is None:
                        return Colouring(m, K, o)
                    if r == 0:
                        continue       # the trivial group colours every bonded pair alike
                    oa = np.asarray(o, np.int64)
                    v = (Kd + oa[alpha2] - oa[alpha]) % mvec[None, :]
                    if not np.any(np.all(v == 0, axis=1)):
                        return Colouring(m, K, o)
    return None


def colour_grid(col: Colouring, b: int, box: tuple[int, ...]) -> np.ndarray:
    """Colour of every site of the box, shape (b, L3, L2, L1), int64."""
    D = len(box)
    L1, L2 = int(box[0]), int(box[1])
    L3 = int(box[2]) if D == 3 else 1
    z, y, x = np.meshgrid(np.arange(L3), np.arange(L2), np.arange(L1), indexing="ij")
    cell = [x, y, z][:D]
    out = np.zeros((b, L3, L2, L1), np.int64)
    for a in range(b):
        idx = np.zeros((L3, L2, L1), np.int64)
        for r, mr in enumerate(col.m):
            g = sum(int(col.K[r][k]) * cell[k] for k in range(D)) + int(col.o[a][r])
            idx = idx * mr + np.mod(g, mr)
        out[a] = idx
    return out


def verify_on_box(col: Colouring, crystal: Crystal, box: tuple[int, ...], periodic: tuple[bool, ...]) -> bool:
    """Colour every site of the box and test every bond explicitly, seam included."""
    D = crystal.D
    cm = colour_grid(col, crystal.b, box)
    L = [int(v) for v in box] + [1] * (3 - D)
    for bd in crystal.bonds:
        d = list(bd.delta) + [0] * (3 - D)
        src = cm[bd.alpha]
        dst = cm[bd.alpha2]
        # neighbour of (z, y, x) is (z + dz, y + dy, x + dx): index arrays per axis
        idx = []
        valid = np.ones(src.shape, bool)
        for k, axis in ((2, 0), (1, 1), (0, 2)):
            coord = np.arange(L[k]) + d[k]
            if k < D and periodic[k]:
                coord = np.mod(coord, L[k])
                ok = np.ones(L[k], bool)
            else:
                ok = (coord >= 0) & (coord < L[k])
                coord = np.clip(coord, 0, L[k] - 1)
            shp = [1, 1, 1]
            shp[axis] = L[k]
            idx.append(coord.reshape(shp))
            valid &= ok.reshape(shp)
        nb = dst[idx[0], idx[1], idx[2]]
        if np.any(valid & (nb == src)):
            return False
    return True


def clique_number(crystal: Crystal, window: int = 2) -> int:
    """Largest clique of the infinite bond graph (Bron-Kerbosch with pivot-free pruning on a window around
    one site of each sublattice; every clique can be translated to contain such a site)."""
    D = crystal.D
    reach = max(max(abs(x) for x in bd.delta) for bd in crystal.bonds)
    w = max(window, reach)
    nodes = [(c, a) for c in itertools.product(range(-w, w + 1), repeat=D) for a in range(crystal.b)]
    index = {nd: i for i, nd in enumerate(nodes)}
    nbrs: list[set[int]] = [set() for _ in nodes]
    for i, (c, a) in enumerate(nodes):
        for bd in crystal.bonds:
            if bd.alpha != a:
                continue
            j = index.get((tuple(c[k] + bd.delta[k] for k in range(D)), bd.alpha2))
            if j is not None:
                nbrs[i].add(j)
    best = 1

    def expand(size: int, cand: set[int]) -> None:
        nonlocal best
        if not cand:
            best = max(best, size)
            return
        if size + len(cand) <= best:
            return
        for v in sorted(cand):
            expand(size + 1, cand & nbrs[v])
            cand = cand - {v}
            if size + len(cand) <= best:
                return

    for a in range(crystal.b):
        start = index[(tuple([0] * D), a)]
        expand(1, set(nbrs[start]))
    return best
