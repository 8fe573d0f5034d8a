"""Sampler: the reference's operator class over the MI355X kernels (SCK/sampler.py:39-345).

Same constructor arguments, attributes and methods as the reference Sampler, with spins in the reference layout
S[B, n, b, L3, L2, L1] (packed Ising: int32 words [W, 1, b, L3, L2, L1]). The sweep, the sequential chain, the
synchronous control and the observables run as hipRTC-compiled kernels through libsck; there is no CPU path,
and a Sampler on a non-HIP device raises.

Additions: `proposal="cone"` (fixed half-angle per replica, `set_width`) and `proposal="learned"` (a width table
from sck_amd.proposal) for XY/Heisenberg, and `observe(S)` returning energy and magnetization from one kernel.
Whole sweeps of a system whose per-replica lattice fits in LDS run in one launch of the resident kernel when
the batch fills the GPU (B >= 128) or the lattice is small (at most 8192 sites); `resident=False` keeps the
per-class launches. A replica-major batch whose lattice exceeds the cache runs
in replica blocks, each through all its sweeps before the next (`replica_block`: replicas per block, a
multiple of 64 dividing the batch; 0 turns blocks off; default automatic).
The Philox counter is 64-bit (identical to the reference's int32 counter below 2^31), so plain batches beyond
the reference's limit run; the bitwise kernels keep the int32 counter w * nsite + i and its limit.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from . import hipgen as G
from .system import System

__all__ = ["Sampler", "choose_layout", "COMPACT_MIN", "TILE", "MAX_BLOCK", "PTILE", "WORD", "BWG", "bitwise_check", "p_table",
           "p_fixed"]

COMPACT_MIN = 4
# the resident kernel runs one workgroup per replica, so it needs a batch that fills the compute units, unless
# the lattice is small enough that per-class launches are launch-bound anyway (measured on int8 Ising: at
# B = 8 and 32 per-class launches are 2-3x faster for 32k-site lattices, the resident kernel 2x faster from
# B = 128, and faster at every B for a 4096-site lattice)
RESIDENT_MIN_BATCH = 128
RESIDENT_SMALL_SITES = 8192
TILE = 128
MAX_BLOCK = 1024
PTILE = 64
WORD = 32
BWG = 4


# ------------------------------------------------------------------------------------- bitwise tables
def bitwise_check(sys_: System) -> tuple[float, float, int]:
    """(J, h, Z) of a system the bitwise kernel can run; ValueError otherwise (SCK/bitwise.py:59-71)."""
    if sys_.n != 1:
        raise ValueError("bitwise multi-spin coding is for Ising (n = 1) only")
    if sys_.extra:
        raise ValueError("bitwise multi-spin coding needs isotropic exchange only (no extra bond matrices)")
    if sys_.cubic_k1 is not None:
        raise ValueError("bitwise multi-spin coding does not support anisotropy")
    Js = {float(sys_.J.get(bd.shell, 0.0)) for bd in sys_.crystal.bonds} - {0.0}
    if len(Js) != 1:
        raise ValueError(f"bitwise multi-spin coding needs one uniform coupling J on every bond, got {sorted(Js)}")
    h = float(sys_.h[0]) if sys_.h is not None else 0.0
    Z = max(sum(1 for bd in sys_.crystal.bonds if bd.alpha == a and sys_.J.get(bd.shell, 0.0) != 0.0)
            for a in range(sys_.b))
    return Js.pop(), h, Z


def _p64(sys_: System, beta: np.ndarray) -> np.ndarray:
    J, h, Z = bitwise_check(sys_)
    beta = np.asarray(beta, np.float64).reshape(-1)
    p = np.zeros((beta.size, 2, Z + 1, Z + 1), np.float64)
    for si in (0, 1):
        sigma = 2 * si - 1
        for ze in range(Z + 1):
            k = np.arange(ze + 1)
            x = beta[:, None] * (2.0 * J * (ze - 2 * k) + 2.0 * h * sigma)[None, :]
            if sys_.acceptance == "barker":
                p[:, si, ze, :ze + 1] = 0.5 * (1.0 - np.tanh(0.5 * x))
            else:
                p[:, si, ze, :ze + 1] = np.exp(-np.maximum(x, 0.0))
    return p


def p_table(sys_: System, beta: np.ndarray) -> np.ndarray:
    """Shared-mode acceptance p[word, s, z_eff, k], float32 (SCK/bitwise.py:79-84)."""
    return _p64(sys_, beta).astype(np.float32)


def p_fixed(sys_: System, beta: np.ndarray) -> np.ndarray:
    """Independent mode P = floor(p 2^32) in [0, 2^32], int64 (SCK/bitwise.py:87-93)."""
    return np.clip(np.floor(_p64(sys_, beta) * 2.0 ** 32).astype(np.int64), 0, 2 ** 32)


def choose_layout(system: System, mode: int, B: int, resident: bool = True, layout: str = "auto", tiled=None,
                  tile_dims=None) -> tuple[str, bool]:
    """(layout, tiled) a Sampler runs: "reference" or "replica" storage, and whether whole sweeps use the tiled
    kernel. `resident` says whether the LDS-resident kernel may run (it takes precedence where it fits)."""
    n, b = system.n, system.b
    rm_ok = mode in (G.MODE_INT8, G.MODE_F32)
    resident_ok = resident and G.fits_resident(system, mode)
    tiled_ok = mode == G.MODE_F32 and not resident_ok and G.fits_tiled(system, mode, tile_dims)
    if tiled and not tiled_ok:
        raise ValueError("the tiled sweep needs float32 vector spins on a 2-D lattice larger than a tile and its "
                         "halo (and no LDS-resident lattice)")
    if tiled and layout == "replica":
        raise ValueError("the tiled sweep runs the reference layout")
    if tiled is None:
        # whole sweeps in tiles where the replica-major lattice would exceed the cache, or where the batch is not
        # whole waves of replicas: the lattice is then read about once per sweep instead of once per class
        big = B * n * b * system.L3 * system.L2 * system.L1 * 4 > G.RM_BLOCK_BYTES
        tiled = tiled_ok and layout != "replica" and (big or B % 64 != 0 or layout == "reference")
    if layout == "auto":
        # lanes over replicas need whole waves of replicas; a lattice that fits LDS runs resident instead
        layout = "replica" if (rm_ok and B % 64 == 0 and not resident_ok and not tiled) else "reference"
    return layout, bool(tiled)


def _stream_ptr(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class Sampler:
    def __init__(self, system: System, batch: int, device="cuda", seed: int = 0, compact: bool = True,
                 spin_dtype: torch.dtype | None = None, packed: bool | None = None, bitwise: bool = False,
                 words_per_program: int | None = None, independent: bool = False, bitwise_tile: int | None = None,
                 bitwise_warps: int | None = None, max_block: int = MAX_BLOCK, proposal: str = "uniform",
                 width=None, learned=None, cache_dir=None, resident: bool = True, layout: str = "auto",
                 tile=None, chunks=None, sites_per_block=None, resident_threads=None, replica_block=None,
                 tiled: bool | None = None, tile_dims=None):
        self.sys = system
        self.use_resident = bool(resident)
        self.bitwise = bool(bitwise)
        self.independent = bool(independent)
        if self.independent and not self.bitwise:
            raise ValueError("independent per-replica draws need bitwise=True")
        if self.bitwise:
            _, _, self.zmax = bitwise_check(system)
            if packed is False:
                raise ValueError("bitwise multi-spin coding uses the packed storage")
            packed = True
        self.packed = (system.n == 1 and int(batch) >= WORD) if packed is None else bool(packed)
        if self.packed and system.n != 1:
            raise ValueError("bit packing is for Ising (n = 1) only")
        self.spin_dtype = spin_dtype or (torch.int8 if system.n == 1 else torch.float32)
        if not self.packed and self.spin_dtype not in (torch.int8, torch.float32, torch.float64):
            raise ValueError("spins are stored as int8 (Ising), float32 or float64")
        if self.spin_dtype == torch.int8 and system.n != 1:
            raise ValueError("int8 storage holds Ising spins only")
        self.B = int(batch)
        if self.B < 1:
            raise ValueError("batch must be at least 1")
        n, b, L3, L2, L1 = system.n, system.b, system.L3, system.L2, system.L1
        self.W = -(-self.B // WORD) if self.packed else self.B
        self.Bp = self.W * WORD if self.packed else self.B
        nsite = b * L3 * L2 * L1
        self.use_resident = self.use_resident and (self.B >= RESIDENT_MIN_BATCH or nsite <= RESIDENT_SMALL_SITES)
        if self.bitwise and self.W * nsite >= 2 ** 31 - 1:
            raise ValueError("replica batch too large for the int32 Philox counter")
        gmax = words_per_program or BWG
        if gmax & (gmax - 1):
            raise ValueError("words_per_program must be a power of two")
        self.G = min(gmax, 1 << (self.W - 1).bit_length()) if self.bitwise else 1
        self.ngroups = -(-self.W // self.G)
        if max_block < 16 or max_block & (max_block - 1):
            raise ValueError("max_block must be a power of two >= 16")
        self.nrows = L3 * L2
        P = G.x_period(system.colouring)
        self.stride = P if (compact and P >= COMPACT_MIN) else 1
        self.bx = -(-L1 // self.stride)
        if proposal != "uniform" and n == 1:
            raise ValueError("a cone proposal needs a vector model (n = 2 or 3)")
        self.proposal = proposal
        if self.bitwise:
            self.mode = G.MODE_BITWISE_INDEP if self.independent else G.MODE_BITWISE
        elif self.packed:
            self.mode = G.MODE_PACKED
        else:
            # float64 storage is updated in float32 arithmetic, as the reference loads any storage type
            self.mode = {torch.int8: G.MODE_INT8, torch.float64: G.MODE_F64}.get(self.spin_dtype, G.MODE_F32)
        self.learned_meta = learned.meta() if learned is not None else None
        if layout not in ("auto", "reference", "replica"):
            raise ValueError("layout must be 'auto', 'reference' or 'replica'")
        rm_ok = self.mode in (G.MODE_INT8, G.MODE_F32)
        if layout == "replica" and not rm_ok:
            raise ValueError("the replica-major layout holds unpacked spins only")
        layout, self.use_tiled = choose_layout(system, self.mode, self.B, self.use_resident, layout, tiled, tile_dims)
        self.replica_major = layout == "replica"
        self.layout = layout
        # arguments are validated above, before the device (as the reference does); nothing runs on the host
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("the MI355X backend runs on a HIP device (device='cuda'); it has no CPU path")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.module = G.generate(system, self.mode, self.B, proposal, self.learned_meta,
                                 replica_major=self.replica_major, rm_tile=tile, rm_chunks=chunks,
                                 rm_sites=sites_per_block, resident_threads=resident_threads,
                                 rm_block=replica_block, tiled=self.use_tiled, tile_dims=tile_dims)
        self.src = self.module.src
        dev = self.device
        m = system.site_mask()
        self.mask_np = m
        self.mask = torch.as_tensor(m.astype(np.int8), device=dev)
        k1 = system.cubic_k1 if system.cubic_k1 is not None else np.zeros(m.shape, np.float32)
        self.k1 = torch.as_tensor(np.asarray(k1, np.float32), device=dev)
        self.state = torch.tensor([int(seed), 0], dtype=torch.int64, device=dev)
        self.beta = torch.ones(self.Bp, dtype=torch.float32, device=dev)
        self._beta_np = np.ones(self.Bp, np.float32)
        if self.bitwise:
            self._table = p_fixed if self.independent else p_table
            self.ptab = torch.as_tensor(self._table(system, np.ones(self.W)), device=dev)
            self.table = self.ptab
        elif n == 1:
            self.table = torch.as_tensor(G.ising_table(system, self._beta_np), device=dev).contiguous()
        else:
            self.table = torch.zeros(1, dtype=torch.float32, device=dev)
        if proposal == "cone":
            w = np.full(self.Bp, np.pi, np.float32) if width is None else np.broadcast_to(
                np.asarray(width, np.float32).reshape(-1), (self.Bp,))
            self.width = torch.as_tensor(np.array(w, np.float32), device=dev)
        elif proposal == "learned":
            if learned is None:
                raise ValueError("proposal='learned' needs a trained LearnedWidth (sck_amd.proposal.train_width)")
            self.width = torch.as_tensor(np.ascontiguousarray(learned.table()), device=dev)
            self._learned = learned
        else:
            self.width = torch.zeros(1, dtype=torch.float32, device=dev)
        ku = system.site_ku if system.site_ku is not None else np.zeros(m.shape, np.float32)
        ax = system.site_axis if system.site_axis is not None else np.zeros(m.shape + (n,), np.float32)
        self.site_ku = torch.as_tensor(np.ascontiguousarray(ku, np.float32), device=dev)
        self.site_axis = torch.as_tensor(np.ascontiguousarray(ax, np.float32), device=dev)
        self.colours = system.colour_map()
        self.classes = self.module.classes
        self._mask_f = torch.as_tensor(m, device=dev).to(torch.float32)
        nv = 1 + b * n
        self._obs = torch.zeros(self.B * nv, dtype=torch.float64, device=dev)
        self._obs_scratch = torch.zeros(self.module.obs_rows * self.module.obs_chunks * nv, dtype=torch.float64,
                                        device=dev)
        self._handle = None
        self._create(cache_dir)

    # ------------------------------------------------------------------------------ native handle
    def _create(self, cache_dir) -> None:
        mod = self.module
        names = (C.c_char_p * len(mod.kernels))(*[k.encode() for k in mod.kernels])
        self._keep = [names]
        d = L.Desc()
        d.abi_version = L.SCK_ABI_VERSION
        d.mode = self.mode
        d.n, d.b, d.L1, d.L2, d.L3 = self.sys.n, self.sys.b, self.sys.L1, self.sys.L2, self.sys.L3
        d.batch, d.words, d.obs_chunks = self.B, self.W, mod.obs_chunks
        d.rep_blocks = mod.rep_blocks
        d.obs_blocks = mod.obs_blocks
        d.kernel_names = C.cast(names, C.POINTER(C.c_char_p))
        d.n_kernels = len(mod.kernels)
        for field_, plan in (("sweep", mod.sweep), ("ref", mod.ref), ("sync", mod.sync), ("obs", mod.obs)):
            arr, cnt = L.launches(plan)
            self._keep.append(arr)
            setattr(d, field_, C.cast(arr, C.POINTER(L.LaunchC)))
            setattr(d, "n_" + field_, cnt)
        barr, _ = L.launches([mod.bump])
        self._keep.append(barr)
        d.bump = C.cast(barr, C.POINTER(L.LaunchC))
        if mod.resident is not None and self.use_resident:
            rarr, _ = L.launches([mod.resident])
            self._keep.append(rarr)
            d.resident = C.cast(rarr, C.POINTER(L.LaunchC))
        if mod.tiled is not None:
            tarr, _ = L.launches([mod.tiled])
            farr, _ = L.launches([mod.track_fold])
            self._keep += [tarr, farr]
            d.tiled = C.cast(tarr, C.POINTER(L.LaunchC))
            d.track_fold = C.cast(farr, C.POINTER(L.LaunchC))
        h = C.c_void_p()
        cache = str(cache_dir or L.CACHE).encode()
        with torch.cuda.device(self.device):
            L.check(L.lib().sck_create(C.byref(d), self.src.encode(), cache, C.byref(h)), "sck_create")
        self._handle = h

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                L.lib().sck_destroy(h)
            except Exception:
                pass
            self._handle = None

    def _workspace(self) -> torch.Tensor:
        """The tiled sweep's second spin buffer (the sweeps alternate between it and the spins), allocated on
        first use and kept with the sampler."""
        if getattr(self, "_ws", None) is None:
            s = self.sys
            self._ws = torch.empty(self.B * s.n * s.nsite, dtype=torch.float32, device=self.device)
        return self._ws

    def _args(self, S: torch.Tensor, dst: torch.Tensor | None = None, sweeps: bool = False) -> L.Args:
        """sck_args over S; `sweeps`: for whole coloured sweeps (the tiled sweep's workspace as dst, and the
        running totals of E and M when they are being tracked)."""
        a = L.Args()
        a.spins = S.data_ptr()
        if dst is None and sweeps and self.use_tiled:
            dst = self._workspace()
        a.dst = dst.data_ptr() if dst is not None else None
        if sweeps and getattr(self, "_tracking", False):
            a.track = self._track.data_ptr()
            a.track_parts = self._track_parts.data_ptr()
        a.mask = self.mask.data_ptr()
        a.k1 = self.k1.data_ptr()
        a.beta = self.beta.data_ptr()
        a.table = self.table.data_ptr()
        a.state = self.state.data_ptr()
        a.width = self.width.data_ptr()
        a.site_ku = self.site_ku.data_ptr()
        a.site_axis = self.site_axis.data_ptr()
        a.obs = self._obs.data_ptr()
        a.obs_scratch = self._obs_scratch.data_ptr()
        a.sweep_offset = 0
        a.nrep = self.W if self.mode in (G.MODE_PACKED, G.MODE_BITWISE, G.MODE_BITWISE_INDEP) else self.B
        a.nsweep = 0
        return a

    def _check_spins(self, S: torch.Tensor) -> None:
        s = self.sys
        if self.packed:
            shape, dt = (self.W, 1) + s.shape, torch.int32
        else:
            shape, dt = (self.B, s.n) + s.shape, self.spin_dtype
        if tuple(S.shape) != shape or S.dtype != dt:
            raise ValueError(f"spins must be {dt} of shape {shape}, got {S.dtype} {tuple(S.shape)}")
        if S.device != self.device:
            raise ValueError(f"spins must be on {self.device}")

    def _is_native(self, S: torch.Tensor) -> bool:
        """Spins in the storage order the kernels index: the reference layout, or for replica-major samplers
        the (B, n, b, L3, L2, L1) view of [n][b][L3][L2][L1][B] storage that init_spins returns."""
        if not self.replica_major:
            return S.is_contiguous()
        return S.permute(2, 3, 4, 5, 1, 0).is_contiguous()

    def _native(self, S: torch.Tensor) -> torch.Tensor:
        """S in the kernels' storage order: contiguous, or (replica-major) the (B, n, b, L3, L2, L1) view of
        contiguous [b][L3][L2][L1][n][B] storage."""
        if self._is_native(S):
            return S
        if not self.replica_major:
            return S.contiguous()
        return S.permute(2, 3, 4, 5, 1, 0).contiguous().permute(5, 4, 0, 1, 2, 3)

    def _run(self, S: torch.Tensor, fn) -> torch.Tensor:
        """fn(native spins) in place on S, through a native-order copy when S is in the other order."""
        self._check_spins(S)
        if self._is_native(S):
            fn(S)
            return S
        T = self._native(S)
        fn(T)
        S.copy_(T)
        return S

    # ------------------------------------------------------------------------------------ state
    def set_temperature(self, T):
        """beta = 1/T per replica, written in place (a captured graph reads the buffer by address)
        (SCK/sampler.py:134-146)."""
        T = np.asarray(torch.as_tensor(T, dtype=torch.float32).reshape(-1).cpu().numpy(), np.float32)
        beta = (np.float32(1.0) / T).astype(np.float32)
        beta = np.broadcast_to(beta, (self.B,)).copy() if beta.size == 1 else beta
        if beta.size != self.B:
            raise ValueError(f"need 1 or {self.B} temperatures, got {beta.size}")
        if self.Bp > self.B:
            beta = np.concatenate([beta, np.full(self.Bp - self.B, beta[-1], np.float32)])
        if self.bitwise:
            bwd = beta.reshape(self.W, WORD)
            if not np.all(bwd == bwd[:, :1]):
                raise ValueError("bitwise multi-spin coding needs one temperature per word of 32 replicas")
            self.ptab.copy_(torch.as_tensor(self._table(self.sys, bwd[:, 0].astype(np.float64)), device=self.device))
        elif self.sys.n == 1:
            self.table.copy_(torch.as_tensor(G.ising_table(self.sys, beta), device=self.device))
        self._beta_np = beta
        self.beta.copy_(torch.as_tensor(beta, device=self.device))

    def set_width(self, width) -> None:
        """Cone half-angle per replica (proposal='cone'), in place."""
        if self.proposal != "cone":
            raise ValueError("set_width needs proposal='cone'")
        w = np.broadcast_to(np.asarray(width, np.float32).reshape(-1), (self.Bp,)) if np.size(width) == 1 else \
            np.asarray(width, np.float32).reshape(-1)
        if w.size == self.B and self.Bp > self.B:
            w = np.concatenate([w, np.full(self.Bp - self.B, w[-1], np.float32)])
        if np.any(w <= 0) or np.any(w > np.pi + 1e-6):
            raise ValueError("cone half-angles must lie in (0, pi]")
        self.width.copy_(torch.as_tensor(np.array(w, np.float32), device=self.device))

    def set_seed(self, seed: int, sweep: int = 0):
        self.state[0] = int(seed)
        self.state[1] = int(sweep)

    def init_spins(self, kind="random", seed: int = 0) -> torch.Tensor:
        """Initial spins drawn with the torch CPU generator as the reference draws them (SCK/sampler.py:152-167)."""
        s = self.sys
        g = torch.Generator(device="cpu").manual_seed(int(seed))
        shape = (self.B, s.n) + s.shape
        if kind == "ordered":
            x = torch.zeros(shape)
            x[:, 0] = 1.0
        elif s.n == 1:
            x = torch.randint(0, 2, shape, generator=g).float() * 2 - 1
        else:
            x = torch.randn(shape, generator=g)
            x = x / x.norm(dim=1, keepdim=True)
        x = x.to(self.device) * self._mask_f
        if self.packed:
            return self.pack(x)
        return self._native(x.to(self.spin_dtype).contiguous())

    def pack(self, x: torch.Tensor) -> torch.Tensor:
        """(B, 1, b, L3, L2, L1) spins of +-1 -> (W, 1, b, L3, L2, L1) int32 words (SCK/sampler.py:169-178)."""
        bits = (x > 0).to(torch.int64)
        if self.Bp > self.B:
            bits = torch.cat([bits, bits.new_zeros((self.Bp - self.B,) + tuple(bits.shape[1:]))])
        bits = bits.reshape((self.W, WORD) + tuple(bits.shape[1:]))
        sh = torch.arange(WORD, device=bits.device, dtype=torch.int64).reshape((1, WORD) + (1,) * (bits.dim() - 2))
        word = (bits << sh).sum(dim=1)
        word = torch.where(word >= 2 ** 31, word - 2 ** 32, word)
        return word.to(torch.int32).contiguous()

    def unpack(self, S: torch.Tensor) -> torch.Tensor:
        """Packed words -> (B, 1, b, L3, L2, L1) int8 +-1, 0 on masked sites (SCK/sampler.py:180-188)."""
        if not self.packed or S.dtype != torch.int32:
            return S
        w = S.to(torch.int64) & 0xFFFFFFFF
        sh = torch.arange(WORD, device=S.device, dtype=torch.int64).reshape((1, WORD) + (1,) * (S.dim() - 1))
        bits = (w.unsqueeze(1) >> sh) & 1
        x = (bits * 2 - 1).reshape((self.Bp,) + tuple(S.shape[1:]))[: self.B]
        return (x * self._mask_f.to(torch.int64)).to(torch.int8)

    # ----------------------------------------------------------------------------------- sweeps
    def sweep(self, S: torch.Tensor, n: int = 1, classes: tuple[int, int] | None = None) -> torch.Tensor:
        """n coloured sweeps in place; advances the sweep counter (SCK/sampler.py:234-240). `classes=(lo, hi)`
        restricts each sweep to class launches [lo, hi) (per-class checks)."""
        lo, hi = classes if classes is not None else (0, -1)

        def run(T):
            a = self._args(T, sweeps=True)
            with torch.cuda.device(self.device):
                L.check(L.lib().sck_sweep(self._handle, C.byref(a), int(n), int(lo), int(hi),
                                          C.c_void_p(_stream_ptr(self.device))), "sweep")
        return self._run(S, run)

    def sweep_timed(self, S: torch.Tensor, n: int) -> tuple[float, int]:
        """n whole sweeps in place, as sweep(), with HIP events around every sweep-kernel launch; returns the
        summed kernel milliseconds and the number of launches (the counter bump is not counted). Synchronizes."""
        if n <= 0:
            raise ValueError("n must be positive")
        self._check_spins(S)
        if not self._is_native(S):
            raise ValueError("timed sweeps need spins in the sampler's storage order (as init_spins returns them)")
        a = self._args(S, sweeps=True)
        ms, nl = C.c_float(), C.c_int32()
        with torch.cuda.device(self.device):
            L.check(L.lib().sck_sweep_timed(self._handle, C.byref(a), int(n), C.c_void_p(_stream_ptr(self.device)),
                                            C.byref(ms), C.byref(nl)), "timed sweep")
        return float(ms.value), int(nl.value)

    def graph(self, S: torch.Tensor, k: int, observe: bool = False):
        """Capture k sweeps (and, with observe=True, one observation after them) in a HIP graph; returns its
        replay (SCK/sampler.py:242-269). The replay keeps the sampler and S alive."""
        self._check_spins(S)
        if not self._is_native(S):
            raise ValueError("graph capture needs spins in the sampler's storage order (as init_spins returns them)")
        return self._capture(S, None, k, L.CAPTURE_OBSERVE if observe else 0)

    def graph_pair(self, S: torch.Tensor, k: int, observe: bool = False):
        """Two replays for measuring after every k sweeps (k odd) with the tiled sweep and no copies: the first
        sweeps S into the sampler's workspace (and observes it there), the second sweeps the workspace back
        into S. Replayed in turn they run the chain graph() runs; after an odd number of replays the state is in
        the workspace (returned third, with S's shape). For other kernels, or even k, both replays are graph()."""
        self._check_spins(S)
        if not self._is_native(S):
            raise ValueError("graph capture needs spins in the sampler's storage order (as init_spins returns them)")
        flags = L.CAPTURE_OBSERVE if observe else 0
        if not (self.use_tiled and k % 2 == 1):
            g = self._capture(S, None, k, flags)
            return g, g, S
        W = self._workspace().view_as(S)
        there = self._capture(S, W, k, flags | L.CAPTURE_IN_DST)
        back = self._capture(W, S, k, flags | L.CAPTURE_IN_DST)
        return there, back, W

    def _capture(self, S: torch.Tensor, dst, k: int, flags: int):
        a = self._args(S, dst=dst, sweeps=True)
        g = C.c_void_p()
        with torch.cuda.device(self.device):
            torch.cuda.current_stream(self.device).synchronize()
            L.check(L.lib().sck_graph_capture(self._handle, C.byref(a), int(k), int(flags), C.byref(g)),
                    "graph capture")
        dev = self.device
        lib = L.lib()

        def replay():
            with torch.cuda.device(dev):
                L.check(lib.sck_graph_replay(g, C.c_void_p(_stream_ptr(dev))), "graph replay")

        def time_ms(reps: int) -> float:
            ms = C.c_float()
            with torch.cuda.device(dev):
                L.check(lib.sck_time_replays(g, int(reps), C.c_void_p(_stream_ptr(dev)), C.byref(ms)), "timing")
            return float(ms.value)

        class _Keep:
            def __init__(self, handle):
                self.handle = handle

            def __del__(self):
                try:
                    lib.sck_graph_destroy(self.handle)
                except Exception:
                    pass

        replay.keep = (self, S, dst, _Keep(g))
        replay.time_ms = time_ms
        return replay

    def sync_sweep(self, S: torch.Tensor, n: int = 1) -> torch.Tensor:
        """The synchronous control: every site updated from the old state (SCK/sampler.py:271-279)."""
        def run(T):
            dst = torch.empty_like(T)
            a = self._args(T, dst)
            with torch.cuda.device(self.device):
                L.check(L.lib().sck_sync_sweep(self._handle, C.byref(a), int(n), T.numel() * T.element_size(),
                                               C.c_void_p(_stream_ptr(self.device))), "sync sweep")
        return self._run(S, run)

    def reference_sweep(self, S: torch.Tensor, n: int = 1) -> torch.Tensor:
        """Sequential reference chain: sites one at a time in class order, raster order inside a class, with
        the same update and counters (SCK/sampler.py:281-293)."""
        def run(T):
            a = self._args(T)
            with torch.cuda.device(self.device):
                L.check(L.lib().sck_reference_sweep(self._handle, C.byref(a), int(n),
                                                    C.c_void_p(_stream_ptr(self.device))), "reference sweep")
        return self._run(S, run)

    # ------------------------------------------------------------------------------ observables
    def observe(self, S: torch.Tensor):
        """(E (B,) float64, M (B, b, n) float64) from the FP64 reduction kernels."""
        self._check_spins(S)
        T = self._native(S)          # held until the kernels are queued on this stream
        a = self._args(T)
        with torch.cuda.device(self.device):
            L.check(L.lib().sck_observe(self._handle, C.byref(a), C.c_void_p(_stream_ptr(self.device))), "observe")
        B, b, n = self.B, self.sys.b, self.sys.n
        return self._obs[:B].clone(), self._obs[B:].reshape(B, b, n).clone()

    def track(self, S: torch.Tensor):
        """Start (or re-anchor) float64 running totals of the energy and the per-sublattice magnetization: they
        are set from one observation of S now, and every later whole sweep of the tiled kernel (sweep(), graph
        replays) adds the energy and spin changes of its accepted moves, accumulated per (replica, tile) in the
        sweep kernel itself. Between re-anchors they differ from observe() only by the fp32 rounding of each
        move's dE. Needs the tiled sweep (2-D vector lattices in the reference layout); returns tracked()."""
        if not self.use_tiled or self.module.track_fold is None:
            raise ValueError("running totals of E and M need the tiled sweep (2-D vector lattices)")
        if getattr(self, "_track", None) is None:
            nv = 1 + self.sys.b * self.sys.n
            self._track = torch.zeros(self.B * nv, dtype=torch.float64, device=self.device)
            self._track_parts = torch.zeros(self.module.track_parts, dtype=torch.float64, device=self.device)
        self.observe(S)
        self._track.copy_(self._obs)
        self._tracking = True
        return self.tracked()

    def tracked(self):
        """Views (E (B,), M (B, b, n)) of the running totals started by track()."""
        if not getattr(self, "_tracking", False):
            raise ValueError("no running totals: call track(S) first")
        B, b, n = self.B, self.sys.b, self.sys.n
        return self._track[:B], self._track[B:].reshape(B, b, n)

    def untrack(self) -> None:
        """Stop updating the running totals (sweeps no longer accumulate them)."""
        self._tracking = False

    def observed(self):
        """Views (E (B,), M (B, b, n)) of the float64 observation buffers: the last observe(), or the
        observation a graph captured with observe=True wrote on its last replay (read after the replay's
        stream has reached it). They are overwritten by the next observation."""
        B, b, n = self.B, self.sys.b, self.sys.n
        return self._obs[:B], self._obs[B:].reshape(B, b, n)

    def energy(self, S: torch.Tensor) -> torch.Tensor:
        """Total energy per replica, float64 (SCK/sampler.py:326-341)."""
        if self.packed and S.dtype != torch.int32:
            return self._energy_torch(S)
        return self.observe(S)[0]

    def magnetization(self, S: torch.Tensor) -> torch.Tensor:
        """Per-sublattice magnetization (B, b, n), float64 (SCK/sampler.py:343-345)."""
        if self.packed and S.dtype != torch.int32:
            return S.double().sum(dim=(3, 4, 5)).transpose(1, 2)
        return self.observe(S)[1]

    def local_field(self, S: torch.Tensor) -> torch.Tensor:
        """Bond part of the local field of every site, float64 (the quantity of SCK/sampler.py:296-324), computed
        on the host side of the kernels as a check: each bond with a non-zero matrix gathers its neighbour spins
        through a flat cell index (zero weight across an open face) and adds its matrix times them."""
        s = self.sys
        x = self.unpack(S).double()
        shape = x.shape                                  # (B, n, b, L3, L2, L1)
        flat = x.reshape(shape[0], s.n, s.b, -1)
        v = torch.zeros_like(flat)
        for i, bd in enumerate(s.crystal.bonds):
            M = s.bond_matrix(i)
            if not np.any(M):
                continue
            idx, weight = self._bond_gather(bd)
            nb = flat[:, :, bd.alpha2].index_select(2, torch.as_tensor(idx, device=x.device))
            nb = nb * torch.as_tensor(weight, device=x.device)
            v[:, :, bd.alpha] += torch.as_tensor(M, dtype=torch.float64, device=x.device) @ nb
        return v.reshape(shape)

    def _bond_gather(self, bd) -> tuple[np.ndarray, np.ndarray]:
        """Flat cell index (z, y, x raster) of every cell's neighbour along bond bd, and its weight: 1 inside the
        box or across a periodic face, 0 across an open face (the index is then clamped into the box)."""
        s = self.sys
        dims = (s.L1, s.L2, s.L3)                        # lattice axis k -> extent
        coords = np.indices((s.L3, s.L2, s.L1)).reshape(3, -1)[::-1].astype(np.int64)   # rows x, y, z
        weight = np.ones(coords.shape[1])
        for k in range(s.crystal.D):
            c = coords[k] + int(bd.delta[k])
            if s.periodic[k]:
                c %= dims[k]
            else:
                weight *= (c >= 0) & (c < dims[k])
                np.clip(c, 0, dims[k] - 1, out=c)
            coords[k] = c
        return (coords[2] * s.L2 + coords[1]) * s.L1 + coords[0], weight

    def _energy_torch(self, x: torch.Tensor) -> torch.Tensor:
        """Energy of unpacked Ising spins: minus half of s . v over the bonds, minus h . s."""
        s = self.sys
        x64 = x.double().reshape(x.shape[0], -1)
        E = -0.5 * (x64 * self.local_field(x).reshape(x.shape[0], -1)).sum(dim=1)
        if s.h is not None:
            E = E - float(s.h[0]) * x64.sum(dim=1)
        return E
