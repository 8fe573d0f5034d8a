"""Crystals (Bravais vectors plus a basis) and their directed bond tables.

Mirrors the crystal model of the reference (SCK/crystals.py:17-109). A site is (c, alpha): an integer cell
coordinate c in Z^D and a basis index alpha. A directed bond (alpha, alpha2, delta, shell) joins (c, alpha)
to (c + delta, alpha2). The table order is fixed, because it is the accumulation order of the local field in
every kernel and in the oracle: shell, then alpha, then alpha2, then delta lexicographic over
[-reach, reach]^D with the last axis fastest (SCK/crystals.py:62-79).

The candidate displacements are evaluated as one numpy array per (alpha, alpha2) pair instead of one
Python iteration per delta.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

__all__ = ["Bond", "Crystal", "shell_distances", "build_bonds", "make", "coordination", "KNOWN"]


@dataclass(frozen=True)
class Bond:
    """One directed bond-table entry: (c, alpha) -> (c + delta, alpha2) on neighbour shell `shell`."""

    alpha: int
    alpha2: int
    delta: tuple[int, ...]
    shell: int


@dataclass
class Crystal:
    name: str
    vectors: np.ndarray          # (D, D); row k is the lattice vector a_k in Cartesian units
    basis: np.ndarray            # (b, D) basis offsets in cell (fractional) coordinates
    shells: tuple[int, ...] = (1,)
    bonds: list[Bond] = field(default_factory=list)

    @property
    def D(self) -> int:
        return int(self.vectors.shape[0])

    @property
    def b(self) -> int:
        return int(self.basis.shape[0])

    def position(self, c, alpha) -> np.ndarray:
        return (np.asarray(c, dtype=float) + self.basis[alpha]) @ self.vectors


def _deltas(D: int, reach: int) -> np.ndarray:
    """Every integer displacement in [-reach, reach]^D, lexicographic with the last axis fastest."""
    axes = [np.arange(-reach, reach + 1)] * D
    grid = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.reshape(-1) for g in grid], axis=1)


def _pair_norms(vectors: np.ndarray, basis: np.ndarray, a: int, a2: int, deltas: np.ndarray) -> np.ndarray:
    r = (deltas.astype(float) + (basis[a2] - basis[a])[None, :]) @ vectors
    return np.linalg.norm(r, axis=1)


def shell_distances(vectors: np.ndarray, basis: np.ndarray, nshell: int, reach: int = 3) -> list[float]:
    """The first `nshell` distinct non-zero neighbour distances over every basis pair, rounded to 1e-9
    (SCK/crystals.py:48-59)."""
    deltas = _deltas(vectors.shape[0], reach)
    found: set[float] = set()
    for a in range(basis.shape[0]):
        for a2 in range(basis.shape[0]):
            d = _pair_norms(vectors, basis, a, a2, deltas)
            found.update(round(float(x), 9) for x in d[d > 1e-9])
    return sorted(found)[:nshell]


def build_bonds(vectors: np.ndarray, basis: np.ndarray, shells: tuple[int, ...], reach: int = 3) -> list[Bond]:
    """Directed bond table of the requested (1-based) shells in the documented order."""
    dist = shell_distances(vectors, basis, max(shells), reach)
    deltas = _deltas(vectors.shape[0], reach)
    out: list[Bond] = []
    for s in sorted(shells):
        target = dist[s - 1]
        for a in range(basis.shape[0]):
            for a2 in range(basis.shape[0]):
                hit = np.abs(_pair_norms(vectors, basis, a, a2, deltas) - target) < 1e-7
                for row in deltas[hit]:
                    out.append(Bond(a, a2, tuple(int(v) for v in row), s))
    return out


_S3 = np.sqrt(3.0)
_TRI = np.array([[1.0, 0.0], [0.5, _S3 / 2]])
#: name -> (Bravais vectors, basis in cell coordinates); bcc and fcc use conventional cubic cells
KNOWN = {
    "square": (np.eye(2), np.zeros((1, 2))),
    "triangular": (_TRI, np.zeros((1, 2))),
    "honeycomb": (_TRI, np.array([[0.0, 0.0], [1 / 3, 1 / 3]])),
    "kagome": (_TRI, np.array([[0.0, 0.0], [0.5, 0.0], [0.0, 0.5]])),
    "sc": (np.eye(3), np.zeros((1, 3))),
    "bcc": (np.eye(3), np.array([[0, 0, 0], [0.5, 0.5, 0.5]])),
    "fcc": (np.eye(3), np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])),
}


def make(name: str, shells: tuple[int, ...] = (1,)) -> Crystal:
    """A named crystal with the bonds of `shells` (SCK/crystals.py:82-101)."""
    if name not in KNOWN:
        raise KeyError(f"unknown crystal {name!r}; known: {sorted(KNOWN)}")
    vec, basis = (np.asarray(x, float) for x in KNOWN[name])
    shells = tuple(int(s) for s in shells)
    reach = 3 if max(shells) <= 2 else 4
    return Crystal(name, vec, basis, shells, build_bonds(vec, basis, shells, reach))


def coordination(crystal: Crystal) -> dict[int, int]:
    """Number of bonds leaving each basis site."""
    out = {a: 0 for a in range(crystal.b)}
    for bd in crystal.bonds:
        out[bd.alpha] += 1
    return out
