"""The binding a maintainer adds to the reference package to run its Sampler on an MI355X (SCK/mi355x.py).

It replaces `generator.compile_system` (SCK/generator.py:322-337) and the per-class Triton launch loop of
`Sampler.sweep` (SCK/sampler.py:226-240) with libsck's C-ABI (include/sck.h). It reads only what the reference's
own Sampler holds (SCK/sampler.py:40-131): `sys, packed, bitwise, independent, spin_dtype, B, W, Bp, mask, k1, beta,
state` and, for bitwise storage, `ptab`. Its System may carry only the reference's fields (SCK/system.py:19-34).

What the reference keeps no counterpart of, the binding owns:
  * the Ising acceptance table. The reference kernel evaluates exp in place (SCK/generator.py:303-309); the
    MI355X kernels read exp(-beta dE) per replica from a table built with numpy's float32 exp. The binding
    builds it from `sampler.beta` and rewrites it in place when the temperature changes (`set_temperature`,
    one call the reference's `Sampler.set_temperature` makes after updating beta, as it rewrites `ptab`);
  * zero-filled buffers for the terms the reference's System cannot hold (proposal width, per-site uniaxial
    strength and axis), and the float64 observable buffers, so every pointer of sck_args is valid.

Edits to the reference, one line each:
    Sampler.__init__:        self._mi355x = MI355XKernel(self)            (in place of compile_system)
    Sampler.set_temperature: self._mi355x.set_temperature(self)            (after beta / ptab are written)
    Sampler.sweep:           self._mi355x.sweep(self, S, n)                (in place of the class loop)
    Sampler.reference_sweep, sync_sweep, graph, energy, magnetization: the methods of the same names below.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from . import hipgen
from .system import as_system

__all__ = ["MI355XKernel"]


def _stream(device) -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class MI355XKernel:
    """libsck module of one reference Sampler, driven the way that Sampler drives its Triton kernel."""

    def __init__(self, sampler, cache_dir=None):
        self.sys = as_system(sampler.sys)
        dev = torch.device(sampler.device)
        if dev.type != "cuda":
            raise ValueError("the MI355X backend runs on a HIP device (device='cuda')")
        self.device = dev if dev.index is not None else torch.device("cuda", torch.cuda.current_device())
        s = self.sys
        if sampler.bitwise:
            self.mode = hipgen.MODE_BITWISE_INDEP if sampler.independent else hipgen.MODE_BITWISE
        elif sampler.packed:
            self.mode = hipgen.MODE_PACKED
        else:
            # the reference's storage dtype (SCK/sampler.py:66): int8 Ising, float32, or float64 (updated in
            # float32 arithmetic, as the reference loads any storage type)
            dt = getattr(sampler, "spin_dtype", torch.int8 if s.n == 1 else torch.float32)
            self.mode = {torch.int8: hipgen.MODE_INT8, torch.float64: hipgen.MODE_F64}.get(dt, hipgen.MODE_F32)
        self.B, self.W = int(sampler.B), int(sampler.W)
        self.packed = self.mode in (hipgen.MODE_PACKED, hipgen.MODE_BITWISE, hipgen.MODE_BITWISE_INDEP)
        # the reference layout [B, n, b, L3, L2, L1]: the reference's tensors are used as they are
        self.mod = hipgen.generate(s, self.mode, self.B)
        dev = self.device
        z = np.zeros(s.shape, np.float32)
        self.site_ku = torch.as_tensor(z, device=dev)
        self.site_axis = torch.zeros(s.shape + (s.n,), dtype=torch.float32, device=dev)
        self.width = torch.zeros(1, dtype=torch.float32, device=dev)
        nv = 1 + s.b * s.n
        self.obs = torch.zeros(self.B * nv, dtype=torch.float64, device=dev)
        self.obs_scratch = torch.zeros(self.mod.obs_rows * self.mod.obs_chunks * nv, dtype=torch.float64, device=dev)
        self.table = None
        if s.n == 1 and not sampler.bitwise:
            self.table = torch.zeros((int(sampler.Bp), self.mod.table_entries), dtype=torch.float32, device=dev)
            self.set_temperature(sampler)
        self.h = C.c_void_p()
        self._create(cache_dir)

    def _create(self, cache_dir) -> None:
        mod, s = self.mod, self.sys
        names = (C.c_char_p * len(mod.kernels))(*[k.encode() for k in mod.kernels])
        self.keep = [names]
        d = _lib.Desc()
        d.abi_version = _lib.SCK_ABI_VERSION
        d.mode, d.n, d.b = self.mode, s.n, s.b
        d.L1, d.L2, d.L3 = s.L1, s.L2, s.L3
        d.batch, d.words, d.obs_chunks = self.B, self.W, mod.obs_chunks
        d.rep_blocks = mod.rep_blocks
        d.obs_blocks = mod.obs_blocks
        d.kernel_names, d.n_kernels = C.cast(names, C.POINTER(C.c_char_p)), len(mod.kernels)
        for field, plan in (("sweep", mod.sweep), ("ref", mod.ref), ("sync", mod.sync), ("obs", mod.obs)):
            arr, n = _lib.launches(plan)
            self.keep.append(arr)
            setattr(d, field, C.cast(arr, C.POINTER(_lib.LaunchC)))
            setattr(d, "n_" + field, n)
        barr, _ = _lib.launches([mod.bump])
        self.keep.append(barr)
        d.bump = C.cast(barr, C.POINTER(_lib.LaunchC))
        cache = str(cache_dir or _lib.CACHE).encode()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().sck_create(C.byref(d), mod.src.encode(), cache, C.byref(self.h)), "sck_create")

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                _lib.lib().sck_destroy(h)
            except Exception:
                pass
            self.h = None

    # ------------------------------------------------------------------------------ state
    def set_temperature(self, sampler) -> None:
        """Rewrite the Ising table in place from sampler.beta (a captured graph reads it by address); the
        bitwise tables are the reference's own `ptab`, which its set_temperature already rewrote."""
        if self.table is None:
            return
        beta = sampler.beta.detach().float().cpu().numpy()
        self.table.copy_(torch.as_tensor(hipgen.ising_table(self.sys, beta), device=self.device))

    def args(self, sampler, S: torch.Tensor, dst: torch.Tensor | None = None) -> _lib.Args:
        if not S.is_contiguous():
            raise ValueError("spins must be contiguous in the reference layout")
        a = _lib.Args()
        a.spins = S.data_ptr()
        a.dst = dst.data_ptr() if dst is not None else None
        a.mask, a.k1 = sampler.mask.data_ptr(), sampler.k1.data_ptr()
        a.beta, a.state = sampler.beta.data_ptr(), sampler.state.data_ptr()
        a.table = sampler.ptab.data_ptr() if sampler.bitwise else \
            (self.table.data_ptr() if self.table is not None else self.width.data_ptr())
        a.width, a.site_ku, a.site_axis = self.width.data_ptr(), self.site_ku.data_ptr(), self.site_axis.data_ptr()
        a.obs, a.obs_scratch = self.obs.data_ptr(), self.obs_scratch.data_ptr()
        a.sweep_offset = 0
        a.nrep = self.W if self.packed else self.B
        return a

    # ----------------------------------------------------------------------------- sweeps
    def sweep(self, sampler, S: torch.Tensor, n: int = 1) -> torch.Tensor:
        """n coloured sweeps in place, then state[1] += n (SCK/sampler.py:234-240)."""
        a = self.args(sampler, S)
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().sck_sweep(self.h, C.byref(a), int(n), 0, -1, _stream(self.device)), "sweep")
        return S

    def reference_sweep(self, sampler, S: torch.Tensor, n: int = 1) -> torch.Tensor:
        """The sequential single-site chain with the same counters (SCK/sampler.py:281-293)."""
        a = self.args(sampler, S)
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().sck_reference_sweep(self.h, C.byref(a), int(n), _stream(self.device)),
                       "reference sweep")
        return S

    def sync_sweep(self, sampler, S: torch.Tensor, n: int = 1) -> torch.Tensor:
        """The synchronous control (SCK/sampler.py:271-279)."""
        dst = torch.empty_like(S)
        a = self.args(sampler, S, dst)
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().sck_sync_sweep(self.h, C.byref(a), int(n), S.numel() * S.element_size(),
                                                 _stream(self.device)), "sync sweep")
        return S

    def graph(self, sampler, S: torch.Tensor, k: int):
        """A replay of k captured sweeps (SCK/sampler.py:242-269); it keeps the sampler, S and this module alive."""
        a = self.args(sampler, S)
        g = C.c_void_p()
        with torch.cuda.device(self.device):
            torch.cuda.current_stream(self.device).synchronize()
            _lib.check(_lib.lib().sck_graph_capture(self.h, C.byref(a), int(k), 0, C.byref(g)), "graph capture")
        dev, lib = self.device, _lib.lib()

        def replay():
            with torch.cuda.device(dev):
                _lib.check(lib.sck_graph_replay(g, _stream(dev)), "graph replay")

        class _Graph:
            def __init__(self, handle):
                self.handle = handle

            def __del__(self):
                try:
                    lib.sck_graph_destroy(self.handle)
                except Exception:
                    pass

        replay.keep = (sampler, S, self, _Graph(g))
        return replay

    # ------------------------------------------------------------------------ observables
    def _observe(self, sampler, S: torch.Tensor):
        a = self.args(sampler, S)
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().sck_observe(self.h, C.byref(a), _stream(self.device)), "observe")
        s = self.sys
        return self.obs[:self.B].clone(), self.obs[self.B:].reshape(self.B, s.b, s.n).clone()

    def energy(self, sampler, S: torch.Tensor) -> torch.Tensor:
        """Total energy per replica, float64 (SCK/sampler.py:326-341)."""
        return self._observe(sampler, S)[0]

    def magnetization(self, sampler, S: torch.Tensor) -> torch.Tensor:
        """Per-sublattice magnetization (B, b, n), float64 (SCK/sampler.py:343-345)."""
        return self._observe(sampler, S)[1]
