"""CPU oracle for the sweep path. TEST INFRASTRUCTURE: only tests/, the smoke check and bench.py's
CPU-baseline leg may import it, and only as the checker (the CPU baseline also times it). The product
(sck_amd + libsck.so) never imports it and has no CPU fallback.

`liboracle.so` (oracle.c, built by `make -C oracle`) restates the reference's sequential single-site chains;
this module fills its system descriptor from a System and adds the small numpy restatements the checks need:
the acceptance tables of the bitwise kernels (SCK/bitwise.py:79-110), the replica initialisation of the
reference's Sampler (SCK/sampler.py:152-188) and float32 exp as the reference's CPU chain evaluates it.
"""

from .oracle import *  # noqa: F401,F403
