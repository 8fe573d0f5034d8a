"""MI355X-native backend for the exact multi-colour Metropolis sweep of O(n) spin models.

The host side mirrors the reference package's system model and Sampler API; the sweep, the bitwise
32-replica Ising update and the observable reductions run as HIP kernels generated per system, compiled
with hipRTC for gfx950 and launched through the C-ABI library libsck.so.
"""

from . import colouring, crystals, system  # noqa: F401

__version__ = "0.1.0"

# the short interface of the reference's `sck` module (describe, check, run, load, Results), loaded on first
# use so that importing the package does not import torch: `import sck_amd as sck; sck.check(config)`
_SHORT = ("describe", "check", "run", "load", "Results")


def __getattr__(name):
    if name in _SHORT:
        from . import api
        return getattr(api, name)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
