"""MI355X-native backend for the exact multi-colour Metropolis sweep of O(n) spin models.

The host side mirrors the reference package's system model and Sampler API; the sweep, the bitwise
32-replica Ising update and the observable reductions run as HIP kernels generated per system, compiled
with hipRTC for gfx950 and launched through the C-ABI library libsck.so.
"""

from . import colouring, crystals, system  # noqa: F401

__version__ = "0.1.0"
