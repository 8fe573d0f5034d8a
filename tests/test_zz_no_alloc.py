"""sck_observe_noalloc: the observables with the per-block partial sums in the library's static device array
instead of a caller workspace. Called through the C-ABI and compared with sck_observe on the same spins, on two
calls in a row (the second on a smaller input, so the array holds the first call's larger contents), and refused
with a clear error when the partial sums exceed the array."""

import ctypes as C

import pytest
import torch

from sck_amd import _lib as L
from sck_amd import system as sy
from sck_amd.sampler import Sampler

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _noalloc(A, S):
    """E (B,) and M (B, b, n) from sck_observe_noalloc, written into a fresh output buffer."""
    s = A.sys
    out = torch.full((A.B * (1 + s.b * s.n),), float("nan"), dtype=torch.float64, device=DEV)
    a = A._args(A._native(S))
    a.obs = out.data_ptr()
    a.obs_scratch = None
    L.check(L.lib().sck_observe_noalloc(A._handle, C.byref(a), C.c_void_p(torch.cuda.current_stream().cuda_stream)),
            "observe_noalloc")
    torch.cuda.synchronize()
    return out[:A.B], out[A.B:].reshape(A.B, s.b, s.n)


def _case(s, B, **kw):
    A = Sampler(s, B, DEV, seed=4, **kw)
    A.set_temperature(torch.linspace(0.5, 2.5, B))
    S = A.init_spins(seed=2)
    A.sweep(S, 3)
    nbytes = A.module.obs_rows * A.module.obs_chunks * (1 + s.b * s.n) * 8
    E0, M0 = A.observe(S)                                   # the entry point with a caller workspace
    E1, M1 = _noalloc(A, S)
    assert torch.equal(E0, E1) and torch.equal(M0, M1)
    return nbytes


def test_noalloc_observe_equals_observe_on_two_calls():
    print("[gpu] observe_noalloc: Heisenberg bcc 8^3 B=512, then XY triangular 12^2 B=64", flush=True)
    first = _case(sy.heisenberg("bcc", 8), 512, layout="replica")
    second = _case(sy.xy("triangular", 12), 64)
    assert second < first <= L.OBS_SCRATCH_BYTES


def test_noalloc_observe_refuses_more_than_its_array():
    # Heisenberg fcc (4 sublattices), one unblocked replica-major launch: 393216 replicas x 1 partial x 13 sums
    # x 8 bytes = 40.9 MB
    s = sy.heisenberg("fcc", 2)
    A = Sampler(s, 393216, DEV, seed=1, layout="replica")
    assert A.module.obs_rows == A.B
    need = A.module.obs_rows * A.module.obs_chunks * (1 + s.b * s.n) * 8
    assert need > L.OBS_SCRATCH_BYTES
    S = A.init_spins(seed=1)
    with pytest.raises(ValueError, match="static scratch"):
        _noalloc(A, S)
    E, _ = A.observe(S)                                     # the workspace entry point still serves it
    assert torch.isfinite(E).all()


def test_noalloc_observe_with_replica_blocks():
    """A replica-blocked module (4 blocks of 64 replicas): every block's partial sums reuse the same rows of the
    static array, and the result equals sck_observe's."""
    print("[gpu] observe_noalloc: Heisenberg bcc 8^3 B=256 in 4 replica blocks", flush=True)
    A = Sampler(sy.heisenberg("bcc", 8), 256, DEV, seed=4, layout="replica", replica_block=64)
    assert A.module.obs_blocks == 4 and A.module.obs_rows == 64
    A.set_temperature(torch.linspace(0.5, 2.5, 256))
    S = A.init_spins(seed=2)
    A.sweep(S, 3)
    E0, M0 = A.observe(S)
    E1, M1 = _noalloc(A, S)
    assert torch.equal(E0, E1) and torch.equal(M0, M1)
