"""Application observables on device spins: the Berg-Luescher topological charge of a layer
(SCK/anneal.py:69-79) and the structure factor at a wave vector (SCK/exp_perf.py:1012-1032).

Both read the spins the sweep leaves in HBM (reference layout (B, n, b, L3, L2, L1) or the replica-major view)
and reduce in float64 on the device.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from .system import System

__all__ = ["topological_charge", "structure_factor", "site_positions"]


def _solid_angle(a, b, c):
    """Signed solid angle of the spherical triangle (a, b, c) (Berg and Luescher, Nucl. Phys. B 190 (1981) 412)."""
    num = (a * torch.cross(b, c, dim=-1)).sum(-1)
    den = 1.0 + (a * b).sum(-1) + (b * c).sum(-1) + (c * a).sum(-1)
    return 2.0 * torch.atan2(num, den)


def topological_charge(S: torch.Tensor, s: System, layer: int | None = None) -> torch.Tensor:
    """Q per replica of Heisenberg spins on layer `layer` (default L3 // 2) of sublattice 0: each plaquette
    (x, y), (x+1, y), (x+1, y+1), (x, y+1) with all four corners active contributes the solid angles of its
    two triangles, and Q is their sum over 4 pi. Open plaquettes across a periodic seam are not counted."""
    if s.n != 3:
        raise ValueError("the topological charge needs Heisenberg spins (n = 3)")
    z = s.L3 // 2 if layer is None else int(layer)
    m = torch.as_tensor(s.site_mask()[0, z], device=S.device)
    sp = S[:, :, 0, z].double().permute(0, 2, 3, 1)                 # (B, L2, L1, 3) indexed [y, x]
    s00, s10, s01, s11 = sp[:, :-1, :-1], sp[:, :-1, 1:], sp[:, 1:, :-1], sp[:, 1:, 1:]
    valid = (m[:-1, :-1] & m[:-1, 1:] & m[1:, :-1] & m[1:, 1:]).double()
    w = _solid_angle(s00, s10, s11) + _solid_angle(s00, s11, s01)
    return (w * valid).sum(dim=(1, 2)) / (4.0 * math.pi)


def site_positions(s: System) -> np.ndarray:
    """Cartesian position of every site, (b, L3, L2, L1, 3)."""
    c = s.crystal
    z, y, x = np.meshgrid(np.arange(s.L3), np.arange(s.L2), np.arange(s.L1), indexing="ij")
    cells = np.stack([x, y, z][:c.D], axis=-1).astype(float)
    out = np.zeros((c.b, s.L3, s.L2, s.L1, 3))
    for a in range(c.b):
        out[a, ..., :c.D] = (cells + c.basis[a]) @ c.vectors
    return out


def structure_factor(S: torch.Tensor, s: System, K) -> torch.Tensor:
    """S(K) = |sum_i S_i exp(i K . r_i)|^2 / N per replica over the active sites (N of them)."""
    K = np.zeros(3) if K is None else np.pad(np.asarray(K, float), (0, 3 - len(K)))
    pos = site_positions(s)
    mask = s.site_mask()
    ph = np.exp(1j * (pos @ K)) * mask
    phase = torch.as_tensor(ph, dtype=torch.complex128, device=S.device)               # (b, L3, L2, L1)
    F = (S.to(torch.complex128) * phase).sum(dim=(2, 3, 4, 5))                         # (B, n)
    return (F.abs() ** 2).sum(dim=1) / float(mask.sum())
