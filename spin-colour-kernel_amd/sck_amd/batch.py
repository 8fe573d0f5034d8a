"""Batched uneven systems: many systems of one Hamiltonian structure, each with its own replicas, swept together.

The reference runs one Sampler, one kernel and one launch sequence per state (SCK/simulate.py:485-494); a
ladder of small nanodisks then leaves most of a GPU idle. Here every system of the batch shares the generated
kernels (the structure -- model, crystal, couplings, extra matrices, field, anisotropy form, colouring
classes -- is compiled in; box, mask and per-site constants are runtime data) and one launch covers all
systems and replicas:

  schedule "class"  one persistent launch per colour class and sweep: workgroups stride over a work list of
                    tiles (system, replica, range of that class's compacted active-site list);
  schedule "chain"  one launch for k whole sweeps: a workgroup per (system, replica) chain runs every class of
                    every sweep with barriers inside the workgroup (chains are independent, so no workgroup
                    ever waits for another).

Each system keeps its own Philox seed and counter layout (rep * nsite + own) and its own spin array in the
reference layout [B_s, n, b, L3, L2, L1], so system s of a batch runs exactly the chain that
Sampler(system_s, B_s, seed=seed_s) runs alone, bit for bit, under every schedule. Sites of one class are
independent, so the order of work inside a class launch changes nothing; the class order is the reference's
colour-major order. The autotuner times HIP-graph replays of candidate schedules (work-list order, tile size,
workgroup size, persistent grid, kernel kind) on the batch itself and keeps the fastest.
"""

from __future__ import annotations

import ctypes as C
import json
from dataclasses import asdict, dataclass

import numpy as np
import torch

from . import _lib as L
from . import hipgen as G
from .system import System

__all__ = ["Schedule", "BatchLayout", "BatchSampler", "signature", "candidate_schedules", "XCDS"]

XCDS = 8                      # accelerator complex dies of an MI355X; workgroup i starts on XCD i mod 8


@dataclass(frozen=True)
class Schedule:
    kind: str = "class"       # "class" (a persistent launch per class) or "chain" (a workgroup per chain)
    block: int = 256          # threads per workgroup
    tile: int = 1024          # sites per tile (class kind)
    order: str = "xcd"        # work-list order: "system", "largest" or "xcd"
    grid: int = 0             # persistent workgroups of a class launch (0: one per tile)

    def label(self) -> str:
        if self.kind == "chain":
            return f"chain/{self.block}/{self.order}"
        return f"class/{self.block}/tile{self.tile}/{self.order}/grid{self.grid or 'all'}"


def candidate_schedules() -> list[Schedule]:
    out = []
    for block in (128, 256):
        for mult in (1, 4):
            for order in ("system", "xcd"):
                for grid in (0, 2048):
                    out.append(Schedule("class", block, block * mult, order, grid))
    for block in G.CHAIN_BLOCKS:
        for order in ("largest", "xcd"):
            out.append(Schedule("chain", block, 0, order, 0))
    return out


def signature(s: System) -> tuple:
    """Everything the batch kernels compile in: systems of one batch must agree on all of it."""
    bonds = tuple((bd.alpha, bd.alpha2, tuple(bd.delta), bd.shell) for bd in s.crystal.bonds)
    extra = tuple(sorted((int(i), np.asarray(m, np.float64).tobytes()) for i, m in s.extra.items()))
    return (s.n, s.crystal.name, s.crystal.b, s.crystal.D, bonds, tuple(sorted(s.J.items())), extra,
            None if s.h is None else tuple(float(v) for v in s.h),
            None if s.uniaxial is None else (float(s.uniaxial[0]), tuple(float(v) for v in s.uniaxial[1])),
            float(s.cubic_k2), s.cubic_k1 is None, s.site_ku is None, s.acceptance, tuple(s.periodic),
            tuple(s.classes()))


class BatchLayout:
    """Host side of a batch (no device needed): offsets of every system's spins, sites and replicas, the
    descriptor table, the compacted per-class site lists, and the work lists of a schedule."""

    def __init__(self, systems, reps, seeds, classes, n: int, b: int):
        self.systems, self.reps, self.seeds, self.classes = list(systems), list(reps), list(seeds), list(classes)
        self.n, self.b = n, b
        self.nsys = len(self.systems)
        self.R = sum(self.reps)
        self.rep_off = np.concatenate([[0], np.cumsum(self.reps)[:-1]]).astype(np.int64)
        ns = np.array([s.nsite for s in self.systems], np.int64)
        self.nsite = ns
        self.spin_off = np.concatenate([[0], np.cumsum(np.array(self.reps) * n * ns)[:-1]]).astype(np.int64)
        self.site_off = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
        self.numel = int((np.array(self.reps) * n * ns).sum())
        self._build()

    # ---------------------------------------------------------------------------------- host tables
    def _build(self) -> None:
        """Descriptor table, compacted class site lists and per-site constants (numpy)."""
        ncls = len(self.classes)
        stride = G.batch_stride(ncls)
        desc = np.zeros((self.nsys, stride), np.int64)
        lists = []
        pos = 0
        self.class_count = np.zeros((self.nsys, ncls), np.int64)
        mask_all, k1_all, ku_all, ax_all = [], [], [], []
        for i, s in enumerate(self.systems):
            m = s.site_mask()
            cm = s.colour_map()
            desc[i, :8] = (s.L1, s.L2, s.L3, self.reps[i], self.rep_off[i], self.seeds[i], self.spin_off[i],
                           self.site_off[i])
            for k, (c, a) in enumerate(self.classes):
                z, y, x = np.nonzero((cm[a] == c) & m[a])            # raster order (z, y, x)
                packed = (x.astype(np.uint32) | (y.astype(np.uint32) << G.COORD_BITS)
                          | (z.astype(np.uint32) << (2 * G.COORD_BITS)))
                desc[i, 8 + 2 * k] = pos
                desc[i, 9 + 2 * k] = packed.size
                self.class_count[i, k] = packed.size
                lists.append(packed)
                pos += packed.size
            mask_all.append(m.reshape(-1).astype(np.int8))
            k1 = s.cubic_k1 if s.cubic_k1 is not None else np.zeros(m.shape, np.float32)
            k1_all.append(np.asarray(k1, np.float32).reshape(-1))
            ku = s.site_ku if s.site_ku is not None else np.zeros(m.shape, np.float32)
            ku_all.append(np.asarray(ku, np.float32).reshape(-1))
            ax = s.site_axis if s.site_axis is not None else np.zeros(m.shape + (self.n,), np.float32)
            ax_all.append(np.asarray(ax, np.float32).reshape(-1))
        if pos >= 2 ** 31:
            raise ValueError("too many active sites for the int32 site lists")
        self.desc = desc
        lst = np.concatenate(lists) if lists else np.zeros(0, np.uint32)
        list_off = 2 * (G.BATCH_HDR + desc.size)                 # in uint32 words
        words = np.zeros(list_off + lst.size + (lst.size & 1), np.uint32)
        hdr = np.array([self.nsys, list_off, ncls, 0], np.int64)
        words[:2 * G.BATCH_HDR] = hdr.view(np.uint32)
        words[2 * G.BATCH_HDR:list_off] = desc.reshape(-1).view(np.uint32)
        words[list_off:list_off + lst.size] = lst
        self.lists = lst
        self.aux = words.view(np.int32)
        self.mask = np.concatenate(mask_all)
        self.k1 = np.concatenate(k1_all)
        self.site_ku = np.concatenate(ku_all)
        self.site_axis = np.concatenate(ax_all)
        self.n_active = np.array([s.n_active for s in self.systems], np.int64)
        # chains in global replica order (sequential chain and observables)
        self.chains = np.array([(i, r) for i in range(self.nsys) for r in range(self.reps[i])], np.int32)
        self.chain_list = np.concatenate([self.chains, np.zeros_like(self.chains)], 1).astype(np.int32)

    def chain_order(self, order: str) -> list[list[tuple[int, int]]]:
        """Chains grouped per XCD queue ("xcd": greedy largest-first balance of active sites over the 8
        XCDs) or in one queue ("system": batch order, "largest": most active sites first)."""
        size = {i: int(self.n_active[i]) for i in range(self.nsys)}
        ch = [(int(i), int(r)) for i, r in self.chains]
        if order == "system":
            return [ch]
        ch.sort(key=lambda c: -size[c[0]])
        if order == "largest":
            return [ch]
        if order != "xcd":
            raise ValueError(f"unknown work order {order!r}")
        queues, load = [[] for _ in range(XCDS)], [0] * XCDS
        for c in ch:
            q = int(np.argmin(load))
            queues[q].append(c)
            load[q] += size[c[0]]
        return queues

    @staticmethod
    def interleave(queues: list[list]) -> list:
        """Queue x's j-th item at position XCDS j + x (empty slots as None), so workgroup p -- dispatched to
        XCD p mod 8 -- and every workgroup a multiple of 8 after it work on one queue."""
        if len(queues) == 1:
            return list(queues[0])
        depth = max(len(q) for q in queues)
        out = []
        for j in range(depth):
            for q in queues:
                out.append(q[j] if j < len(q) else None)
        return out

    def class_tiles(self, k: int, sch: Schedule) -> np.ndarray:
        tile = int(sch.tile)
        queues = []
        for q in self.chain_order(sch.order):
            tq = []
            for i, r in q:
                lo = int(self.desc[i, 8 + 2 * k])
                cnt = int(self.desc[i, 9 + 2 * k])
                for st in range(0, cnt, tile):
                    tq.append((i, r, lo + st, lo + min(cnt, st + tile)))
            queues.append(tq)
        items = self.interleave(queues)
        return np.array([t if t is not None else (0, 0, 0, 0) for t in items], np.int32).reshape(-1, 4)

    def chain_tiles(self, sch: Schedule) -> np.ndarray:
        items = self.interleave(self.chain_order(sch.order))
        # an empty slot becomes a chain with no work: replica index -1 is never reached, its sweeps are skipped
        return np.array([(c[0], c[1], 0, 0) if c is not None else (0, 0, 0, -1) for c in items], np.int32)


class BatchSampler:
    """Systems sharing one Hamiltonian structure, `replicas` chains each (an int or one per system)."""

    def __init__(self, systems, replicas=16, device="cuda", seed: int = 0, seeds=None, proposal: str = "uniform",
                 width=None, learned=None, schedule="auto", cache_dir=None):
        self.systems = list(systems)
        if not self.systems:
            raise ValueError("a batch needs at least one system")
        sig = signature(self.systems[0])
        for i, s in enumerate(self.systems[1:], 1):
            if signature(s) != sig:
                raise ValueError(f"system {i} ({s.name}) differs in Hamiltonian structure from system 0 "
                                 f"({self.systems[0].name}); a batch shares one set of kernels")
        s0 = self.systems[0]
        self.n, self.b = s0.n, s0.b
        self.nsys = len(self.systems)
        reps = [int(replicas)] * self.nsys if np.ndim(replicas) == 0 else [int(r) for r in replicas]
        if len(reps) != self.nsys or min(reps) < 1:
            raise ValueError("replicas must be a positive int or one positive int per system")
        self.reps = reps
        self.seeds = [int(seed) + i for i in range(self.nsys)] if seeds is None else [int(v) for v in seeds]
        if len(self.seeds) != self.nsys:
            raise ValueError("need one seed per system")
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("the MI355X backend runs on a HIP device (device='cuda'); it has no CPU path")
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        lim = 1 << G.COORD_BITS
        dmax = max((max(abs(v) for v in bd.delta) for bd in s0.crystal.bonds), default=0)
        for s in self.systems:
            if max(s.L1, s.L2, s.L3) > lim:
                raise ValueError(f"{s.name}: box {s.box} exceeds {lim} sites along an axis")
            for k in range(s.crystal.D):
                if s.periodic[k] and s.box[k] <= dmax:
                    raise ValueError(f"{s.name}: periodic box {s.box} shorter than the longest bond")
        if proposal not in ("uniform", "cone", "learned"):
            raise ValueError("proposal must be 'uniform', 'cone' or 'learned'")
        if proposal != "uniform" and self.n == 1:
            raise ValueError("a cone proposal needs a vector model (n = 2 or 3)")
        self.proposal = proposal
        self.mode = G.MODE_INT8 if self.n == 1 else G.MODE_F32
        self.spin_dtype = torch.int8 if self.n == 1 else torch.float32
        ns = [s.nsite for s in self.systems]
        wide = bool(max(r * n_ for r, n_ in zip(reps, ns)) >= 2 ** 31)
        self.learned_meta = learned.meta() if learned is not None else None
        self.module = G.generate_batch(self.systems, self.mode, proposal, self.learned_meta, wide=wide)
        self.classes = self.module.classes
        lay = BatchLayout(self.systems, reps, self.seeds, self.classes, self.n, self.b)
        self.layout = lay
        self.R, self.rep_off, self.spin_off, self.numel = lay.R, lay.rep_off, lay.spin_off, lay.numel
        dev = self.device
        self.aux = torch.as_tensor(lay.aux, device=dev)
        self.mask = torch.as_tensor(lay.mask, device=dev)
        self.k1 = torch.as_tensor(lay.k1, device=dev)
        self.site_ku = torch.as_tensor(lay.site_ku, device=dev)
        self.site_axis = torch.as_tensor(lay.site_axis, device=dev)
        self._chains_dev = torch.as_tensor(lay.chain_list, device=dev)
        self.state = torch.tensor([self.seeds[0], 0], dtype=torch.int64, device=dev)
        self.beta = torch.ones(self.R, dtype=torch.float32, device=dev)
        self._beta_np = np.ones(self.R, np.float32)
        if self.n == 1:
            self.table = torch.as_tensor(G.ising_table(s0, self._beta_np), device=dev).contiguous()
        else:
            self.table = torch.zeros(1, dtype=torch.float32, device=dev)
        if proposal == "cone":
            w = np.full(self.R, np.pi, np.float32) if width is None else np.broadcast_to(
                np.asarray(width, np.float32).reshape(-1), (self.R,))
            self.width = torch.as_tensor(np.array(w, np.float32), device=dev)
        elif proposal == "learned":
            if learned is None:
                raise ValueError("proposal='learned' needs a trained LearnedWidth")
            self.width = torch.as_tensor(np.ascontiguousarray(learned.table()), device=dev)
        else:
            self.width = torch.zeros(1, dtype=torch.float32, device=dev)
        nv = 1 + self.b * self.n
        self._obs = torch.zeros(self.R * nv, dtype=torch.float64, device=dev)
        self._obs_scratch = torch.zeros(1, dtype=torch.float64, device=dev)
        self.cache_dir = cache_dir
        self._handles: dict[Schedule, tuple] = {}
        self.tuning: list[tuple[str, float]] = []
        self.schedule = None
        self.use(Schedule() if schedule in ("auto", None) else schedule)
        self._autotune_pending = schedule == "auto"

    # ------------------------------------------------------------------------------ native handles
    def _handle_for(self, sch: Schedule):
        if sch in self._handles:
            return self._handles[sch]
        if sch.kind not in ("class", "chain"):
            raise ValueError(f"unknown schedule kind {sch.kind!r}")
        if sch.kind == "class" and not (64 <= sch.block <= G.CLASS_BOUND and sch.block % 64 == 0 and sch.tile >= 1):
            raise ValueError(f"class schedules run 64..{G.CLASS_BOUND} threads (a multiple of 64) per workgroup")
        if sch.kind == "chain" and sch.block not in G.CHAIN_BLOCKS:
            raise ValueError(f"chain schedules run {G.CHAIN_BLOCKS} threads per workgroup")
        if sch.grid and sch.grid % XCDS:
            raise ValueError("a persistent grid must be a multiple of 8 workgroups (one per XCD)")
        mod = self.module
        ki = {k: i for i, k in enumerate(mod.kernels)}
        keep = []
        dflt = Schedule() if sch.kind == "chain" else sch
        sweep = []
        for k in range(len(self.classes)):
            t = torch.as_tensor(self.layout.class_tiles(k, dflt), device=self.device)
            keep.append(t)
            nt = int(t.shape[0])
            grid = nt if not dflt.grid else min(dflt.grid, -(-nt // XCDS) * XCDS)
            sweep.append(G.Launch(ki[f"sck_bclass_{k}"], (max(1, grid), 1, 1), (dflt.block, 1, 1),
                                  tiles_ptr=t.data_ptr(), ntile=nt))
        resident = None
        if sch.kind == "chain":
            t = torch.as_tensor(self.layout.chain_tiles(sch), device=self.device)
            keep.append(t)
            resident = G.Launch(ki[f"sck_bchain_{sch.block}"], (int(t.shape[0]), 1, 1), (sch.block, 1, 1),
                                tiles_ptr=t.data_ptr(), ntile=int(t.shape[0]))
        cp = self._chains_dev.data_ptr()
        ref = [G.Launch(ki["sck_bseq"], (-(-self.R // 64), 1, 1), (64, 1, 1), tiles_ptr=cp, ntile=self.R)]
        obs = [G.Launch(ki["sck_bobs"], (self.R, 1, 1), (G.BLOCK, 1, 1), tiles_ptr=cp, ntile=self.R)]
        bump = G.Launch(ki["sck_bump"], (1, 1, 1), (64, 1, 1))
        names = (C.c_char_p * len(mod.kernels))(*[k.encode() for k in mod.kernels])
        keep.append(names)
        d = L.Desc()
        d.abi_version = L.SCK_ABI_VERSION
        d.mode = self.mode
        d.n, d.b = self.n, self.b
        d.L1 = d.L2 = d.L3 = 0
        d.batch, d.words, d.obs_chunks = self.R, self.R, 1
        d.rep_blocks = 1
        d.obs_blocks = 1
        d.kernel_names = C.cast(names, C.POINTER(C.c_char_p))
        d.n_kernels = len(mod.kernels)
        for field_, plan in (("sweep", sweep), ("ref", ref), ("sync", []), ("obs", obs)):
            arr, cnt = L.launches(plan)
            keep.append(arr)
            setattr(d, field_, C.cast(arr, C.POINTER(L.LaunchC)))
            setattr(d, "n_" + field_, cnt)
        barr, _ = L.launches([bump])
        keep.append(barr)
        d.bump = C.cast(barr, C.POINTER(L.LaunchC))
        if resident is not None:
            rarr, _ = L.launches([resident])
            keep.append(rarr)
            d.resident = C.cast(rarr, C.POINTER(L.LaunchC))
        h = C.c_void_p()
        cache = str(self.cache_dir or L.CACHE).encode()
        with torch.cuda.device(self.device):
            L.check(L.lib().sck_create(C.byref(d), mod.src.encode(), cache, C.byref(h)), "sck_create (batch)")
        self._handles[sch] = (h, keep)
        return self._handles[sch]

    def use(self, sch: Schedule) -> None:
        """Make `sch` the schedule of sweep/graph."""
        self._handle_for(sch)
        self.schedule = sch

    @property
    def _handle(self):
        return self._handles[self.schedule][0]

    def __del__(self):
        for h, _ in getattr(self, "_handles", {}).values():
            try:
                if h.value:
                    L.lib().sck_destroy(h)
            except Exception:
                pass
        self._handles = {}

    def _args(self, S: torch.Tensor) -> L.Args:
        a = L.Args()
        a.spins = S.data_ptr()
        a.dst = None
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
        a.nrep = self.R
        a.nsweep = 0
        a.aux = self.aux.data_ptr()
        return a

    def _check(self, S: torch.Tensor) -> None:
        if S.dtype != self.spin_dtype or S.numel() != self.numel or not S.is_contiguous() or S.device != self.device:
            raise ValueError(f"batch spins must be one contiguous {self.spin_dtype} tensor of {self.numel} "
                             f"elements on {self.device} (as init_spins returns)")

    # ------------------------------------------------------------------------------------ state
    def views(self, S: torch.Tensor) -> list[torch.Tensor]:
        """System s's spins (B_s, n, b, L3, L2, L1), a view into the batch tensor."""
        self._check(S)
        out = []
        for i, s in enumerate(self.systems):
            o, cnt = int(self.spin_off[i]), self.reps[i] * self.n * s.nsite
            out.append(S[o:o + cnt].view((self.reps[i], self.n) + s.shape))
        return out

    def init_spins(self, kind="random", seed: int = 0) -> torch.Tensor:
        """System s drawn as Sampler(system_s).init_spins(kind, seed + s) draws it (torch CPU generator)."""
        S = torch.zeros(self.numel, dtype=self.spin_dtype, device=self.device)
        for i, (s, v) in enumerate(zip(self.systems, self.views(S))):
            g = torch.Generator(device="cpu").manual_seed(int(seed) + i)
            shape = (self.reps[i], s.n) + s.shape
            if kind == "ordered":
                x = torch.zeros(shape)
                x[:, 0] = 1.0
            elif s.n == 1:
                x = torch.randint(0, 2, shape, generator=g).float() * 2 - 1
            else:
                x = torch.randn(shape, generator=g)
                x = x / x.norm(dim=1, keepdim=True)
            x = x * torch.as_tensor(s.site_mask(), dtype=torch.float32)
            v.copy_(x.to(self.device).to(self.spin_dtype))
        return S

    def set_temperature(self, T) -> None:
        """beta = 1/T: one value, one per system (scalars or arrays of its replicas), or one per replica."""
        if isinstance(T, (list, tuple)) and len(T) == self.nsys:
            parts = [np.broadcast_to(np.asarray(t, np.float32).reshape(-1), (r,)) for t, r in zip(T, self.reps)]
            Tv = np.concatenate(parts).astype(np.float32)
        else:
            Tv = np.asarray(torch.as_tensor(T, dtype=torch.float32).reshape(-1).cpu().numpy(), np.float32)
            Tv = np.broadcast_to(Tv, (self.R,)).copy() if Tv.size == 1 else Tv
        if Tv.size != self.R:
            raise ValueError(f"need 1, {self.nsys} (per system) or {self.R} temperatures, got {Tv.size}")
        beta = (np.float32(1.0) / Tv).astype(np.float32)
        if self.n == 1:
            self.table.copy_(torch.as_tensor(G.ising_table(self.systems[0], beta), device=self.device))
        self._beta_np = beta
        self.beta.copy_(torch.as_tensor(beta, device=self.device))

    def set_width(self, width) -> None:
        if self.proposal != "cone":
            raise ValueError("set_width needs proposal='cone'")
        w = np.broadcast_to(np.asarray(width, np.float32).reshape(-1), (self.R,))
        if np.any(w <= 0) or np.any(w > np.pi + 1e-6):
            raise ValueError("cone half-angles must lie in (0, pi]")
        self.width.copy_(torch.as_tensor(np.array(w, np.float32), device=self.device))

    # ----------------------------------------------------------------------------------- sweeps
    def sweep(self, S: torch.Tensor, n: int = 1, classes: tuple[int, int] | None = None) -> torch.Tensor:
        """n coloured sweeps of every system in place (under the current schedule)."""
        self._check(S)
        if self._autotune_pending and classes is None and n > 0:
            self.autotune(S)
        lo, hi = classes if classes is not None else (0, -1)
        a = self._args(S)
        with torch.cuda.device(self.device):
            L.check(L.lib().sck_sweep(self._handle, C.byref(a), int(n), int(lo), int(hi),
                                      C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)), "batch sweep")
        return S

    def sweep_timed(self, S: torch.Tensor, n: int) -> tuple[float, int]:
        """n sweeps as sweep(), with HIP events around every sweep-kernel launch; returns the summed kernel
        milliseconds and the number of launches (the counter bump is not counted). Synchronizes."""
        self._check(S)
        if n <= 0:
            raise ValueError("n must be positive")
        a = self._args(S)
        ms, nl = C.c_float(), C.c_int32()
        with torch.cuda.device(self.device):
            L.check(L.lib().sck_sweep_timed(self._handle, C.byref(a), int(n),
                                            C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream),
                                            C.byref(ms), C.byref(nl)), "batch timed sweep")
        return float(ms.value), int(nl.value)

    def reference_sweep(self, S: torch.Tensor, n: int = 1) -> torch.Tensor:
        """Sequential chain of every (system, replica): classes in order, sites in raster order."""
        self._check(S)
        a = self._args(S)
        with torch.cuda.device(self.device):
            L.check(L.lib().sck_reference_sweep(self._handle, C.byref(a), int(n),
                                                C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)),
                    "batch reference sweep")
        return S

    def graph(self, S: torch.Tensor, k: int, observe: bool = False):
        """k sweeps (and optionally one observation) of the whole batch captured in one HIP graph."""
        self._check(S)
        if self._autotune_pending:
            self.autotune(S)
        a = self._args(S)
        g = C.c_void_p()
        dev = self.device
        with torch.cuda.device(dev):
            torch.cuda.current_stream(dev).synchronize()
            L.check(L.lib().sck_graph_capture(self._handle, C.byref(a), int(k), int(bool(observe)), C.byref(g)),
                    "batch graph capture")
        lib = L.lib()

        def replay():
            with torch.cuda.device(dev):
                L.check(lib.sck_graph_replay(g, C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)), "replay")

        def time_ms(reps: int) -> float:
            ms = C.c_float()
            with torch.cuda.device(dev):
                L.check(lib.sck_time_replays(g, int(reps), C.c_void_p(torch.cuda.current_stream(dev).cuda_stream),
                                             C.byref(ms)), "timing")
            return float(ms.value)

        class _Keep:
            def __init__(self, handle):
                self.handle = handle

            def __del__(self):
                try:
                    lib.sck_graph_destroy(self.handle)
                except Exception:
                    pass

        replay.keep = (self, S, self._handles[self.schedule], _Keep(g))
        replay.time_ms = time_ms
        return replay

    # ------------------------------------------------------------------------------- autotuner
    def autotune(self, S: torch.Tensor, candidates=None, sweeps: int = 4, reps: int = 5, log=None) -> Schedule:
        """Time `reps` replays of a graph of `sweeps` sweeps under each candidate schedule on a copy of S
        (the sweep counter is restored afterwards) and keep the fastest. Every schedule gives the same chain,
        so only time differs."""
        self._autotune_pending = False
        cands = list(candidates) if candidates is not None else candidate_schedules()
        work = S.clone()
        state0 = self.state.clone()
        results = []
        for sch in cands:
            self.use(sch)
            rep = self.graph(work, sweeps)
            rep()
            ms = rep.time_ms(reps) / (reps * sweeps)
            results.append((sch, ms))
            if log:
                log(f"  schedule {sch.label():36s} {ms * 1e3:9.1f} us/sweep")
            del rep
        self.state.copy_(state0)
        best = min(results, key=lambda r: r[1])[0]
        self.tuning = [(s.label(), ms) for s, ms in results]
        # keep only the chosen schedule's handle (after the device has drained the tuning graphs)
        torch.cuda.synchronize(self.device)
        for sch in list(self._handles):
            if sch != best:
                h, _ = self._handles.pop(sch)
                L.lib().sck_destroy(h)
        self.use(best)
        return best

    # ------------------------------------------------------------------------------ observables
    def observe(self, S: torch.Tensor):
        """(E (R,) float64, M (R, b, n) float64) of every (system, replica) in global replica order."""
        self._check(S)
        a = self._args(S)
        with torch.cuda.device(self.device):
            L.check(L.lib().sck_observe(self._handle, C.byref(a),
                                        C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)), "observe")
        R, b, n = self.R, self.b, self.n
        return self._obs[:R].clone(), self._obs[R:].reshape(R, b, n).clone()

    def split(self, x: torch.Tensor) -> list[torch.Tensor]:
        """Per-system slices of a per-replica array (global replica order)."""
        return [x[int(o):int(o) + r] for o, r in zip(self.rep_off, self.reps)]

    def schedule_json(self) -> str:
        return json.dumps(dict(schedule=asdict(self.schedule), tuning=self.tuning))
