"""Per-system HIP source generator for gfx950 (the MI355X counterpart of SCK/generator.py and SCK/bitwise.py).

One System gives one HIP translation unit, compiled by libsck with hipRTC. Everything fixed at generation time
is straight-line code: model, bond table grouped by J value, extra (DMI) matrices, field, anisotropies,
colouring and box. Runtime values are the replica index, beta, the acceptance table and the Philox state.

Kernels of a module (all take one `sck_args` by value):

  sck_class_<k>   one (colour, sublattice) class of the coloured sweep, k in sweep order. Lanes cover the class
                  sites of one row after another (the colouring's x-period P gives each row's sites as
                  x0 + P j), one block row per replica, so loads along x stay coalesced in the reference
                  layout S[B, n, b, L3, L2, L1].
                  With the replica-major layout S[b, L3, L2, L1, n, B] the lanes cover replicas of one site
                  instead, so every neighbour load is a contiguous run over replicas.
  sck_tiled       whole sweeps of a 2-D vector lattice in one launch: a workgroup stages a tile and a halo of
                  one bond length per class in LDS, updates the classes in sweep order on a halo that shrinks
                  by one bond per class, and writes the tile's interior (draws depend only on seed, sweep,
                  replica and site, so halo copies repeat their neighbours' updates bit for bit).
  sck_resident    k whole sweeps of one replica per workgroup with the lattice in LDS (small lattices).
  sck_seq         the sequential chain: one thread per replica (per word for packed/bitwise storage) visiting
                  the classes in sweep order and their sites in raster order (SCK/sampler.py:281-293).
  sck_sync_<a>    synchronous control of sublattice a (SCK/sampler.py:271-279).
  sck_obs, sck_obs_final   float64 energy and magnetization: per-block partial sums, then a fixed-order sum.
  sck_track_fold  adds one tiled sweep's per-(replica, tile) changes of E and M to the running totals.
  sck_sample      records (spin, local field, beta, constants) at chosen (replica, site) pairs for training.
  sck_bump        state[1] += sweep_offset.

A batch module (generate_batch) holds the same kernels over many uneven systems: sck_bclass_<k>, sck_bseq,
sck_bobs, and sck_bchain_<t> for persistent launches of several sweeps.

The per-site update follows SCK/generator.py:202-318 operation by operation in fp32 with contraction off, so
the sequential kernel, the class kernels and the CPU oracle agree:

  * Ising (n = 1): the local field only enters through dE = 2 s v, so the kernel sums each J group's
    neighbour spins as integers and reads exp(-beta dE) (Barker: 1 + exp(beta dE)) from a per-replica table
    built on the host with numpy float32 exp, exactly the values the reference's CPU chain computes. The
    draw is skipped where the table decides alone (p >= 1 or p == 0): the Philox stream is counter-based, so
    skipping a draw changes no other draw.
  * XY / Heisenberg: fp32 local field, proposal and dE in the reference's order. cos and sin come from
    sincos_bounded (Cody-Waite reduction and the Cephes single-precision polynomials, error a few ulp for the
    bounded arguments used here) and exp from the hardware base-2 exponential of x log2(e) (exp_fast: relative
    error about |x| 2^-24 plus an ulp, under 3e-6 wherever |beta dE| <= 50). Neither is the reference's numpy
    float32 function, so parity is the decision tolerance: equal decisions except where |beta dE + ln u| < 1e-5.

Philox-4x32-10: key (seed, 0), counter (rep * nsite + own). Below 2^31 this is exactly the reference's int32
counter (SCK/generator.py:275-276). The reference refuses every batch with Bp * nsite >= 2^31 - 1
(SCK/sampler.py:73-74); this build instead carries the counter in 64 bits, low word in c0 and high word in c1,
so larger batches (bitwise Ising sc 128^3 x 1024, XY triangular 2048^2 x 2048) run. That is a deliberate
extension outside the reference's range, pinned by the oracle's restatement of the same rule.
"""

from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, field, replace

import numpy as np

from .system import System, as_system

__all__ = ["Launch", "Module", "generate", "x_period", "ising_layout", "MODE_INT8", "MODE_F32", "MODE_PACKED",
           "MODE_BITWISE", "MODE_BITWISE_INDEP", "MODE_F64", "BLOCK", "BatchModule", "generate_batch", "batch_stride",
           "BATCH_HDR", "CHAIN_BLOCKS", "COORD_BITS"]

MODE_INT8, MODE_F32, MODE_PACKED, MODE_BITWISE, MODE_BITWISE_INDEP, MODE_F64 = range(6)
BLOCK = 256
XCDS = 8                      # XCDs of an MI355X; consecutive workgroups are dispatched to them round-robin
RESIDENT_GROUP = 4            # resident Ising kernel: sites a thread reads before it draws any of them
RM_WAVES = 7                  # occupancy cap of the replica-major class kernels (waves per SIMD)
WORD = 32
GROUP_KEY_STEP = 104729       # independent bitwise mode: key of level group g (SCK/bitwise.py:159-168)
BW_QUEUE_DIV = 4              # independent bitwise mode: the tail queue holds 1/4 of a block's items
BW_WORDS = 2                  # independent bitwise mode: words per thread, sharing the site's index work
LEVELS = 32
RM_BLOCK_BYTES = 224 << 20    # replica-major: largest lattice share run through whole sweeps at a time (cache-resident)
# tiled sweep: tile sizes (cells along x, y) tried in order, the first whose tile and halo fit the LDS budget.
# XY triangular 2048^2 x 2048: 128 x 32 gives 2.11e11 site updates/s, 64 x 64 2.04e11, 128 x 48 1.99e11,
# 96 x 48 1.99e11 (one box, one run each)
TILE_CANDIDATES = ((128, 32), (64, 64), (64, 48), (64, 32), (32, 32), (24, 24))


@dataclass
class Launch:
    kernel: int
    grid: tuple[int, int, int]
    block: tuple[int, int, int] = (BLOCK, 1, 1)
    shmem: int = 0
    tiles_ptr: int = 0        # device work list handed to the kernel (batched systems), 0 for none
    ntile: int = 0
    rep_base: int = 0         # replica-major class launches: first replica of the block the launch covers


@dataclass
class Module:
    src: str
    kernels: list[str]
    sweep: list[Launch]
    ref: list[Launch]
    sync: list[Launch]
    obs: list[Launch]
    bump: Launch
    obs_chunks: int
    classes: list[tuple[int, int]]
    table_entries: int = 0            # Ising: table entries per replica
    ising_layout: list = field(default_factory=list)
    key: str = ""
    resident: Launch | None = None     # whole sweeps of one replica per workgroup with the lattice in LDS
    tiled: Launch | None = None        # one whole sweep per launch, (replica, tile) per workgroup, tiles in LDS
    tile_dims: tuple | None = None     # the tiled kernel's cells per tile (x, y)
    track_fold: Launch | None = None   # adds a tiled sweep's partial sums of E and M to the running totals
    track_parts: int = 0               # doubles of those partial sums
    rep_blocks: int = 1                # the sweep list is this many groups of class launches, one per replica block
    obs_blocks: int = 1                # the observation list is this many (partial sums, fold) groups, one per block
    obs_rows: int = 0                  # replicas whose partial sums the observation scratch holds at once


def _hexf(v: float) -> str:
    """A constant as the fp32 value the reference bakes into its source (SCK/generator.py:44-45), written as an
    exact C99 hexadecimal literal."""
    return f"{float(np.float32(v)).hex()}f"


def a_expr(var: str, stride: int) -> str:
    """var * stride, folded to 0 for a single sublattice."""
    return f"{var} * {stride}"


def _f64(x: float) -> str:
    return float(x).hex()


def x_period(col) -> int:
    """Period along x of a linear colouring: lcm over factors of m_r / gcd(K_r1, m_r) (SCK/generator.py:58-63)."""
    P = 1
    for r, mr in enumerate(col.m):
        P = math.lcm(P, mr // math.gcd(int(col.K[r][0]) % mr or mr, mr))
    return P


def y_period(col) -> int:
    """Period along y of a linear colouring: lcm over factors of m_r / gcd(K_r2, m_r)."""
    P = 1
    for r, mr in enumerate(col.m):
        ky = int(col.K[r][1]) if len(col.K[r]) > 1 else 0
        P = math.lcm(P, mr // math.gcd(ky % mr or mr, mr))
    return P


# --------------------------------------------------------------------------------------- bond plans
def _groups(sys_: System, a: int):
    """Bonds of sublattice a grouped by J value in first-appearance order, and the extras in bond order
    (SCK/generator.py:222-233)."""
    groups: dict[float, list[int]] = {}
    extras: list[int] = []
    for i, bd in enumerate(sys_.crystal.bonds):
        if bd.alpha != a:
            continue
        groups.setdefault(float(sys_.J.get(bd.shell, 0.0)), []).append(i)
        if i in sys_.extra:
            extras.append(i)
    out = []
    for Jv, members in groups.items():
        if Jv == 0.0 and not any(i in sys_.extra for i in members):
            continue
        out.append((Jv, members))
    return out, extras


def ising_layout(sys_: System):
    """Digits of the Ising table index per sublattice: [(kind, size, payload)], kind 'g' (a J group's
    neighbour sum + n), 'e' (an extra bond's neighbour value + 1) or 's' (own spin > 0). Returns
    (per-sublattice digit lists, per-sublattice offsets, entries per replica)."""
    layouts, offsets, total = [], [], 0
    for a in range(sys_.b):
        groups, extras = _groups(sys_, a)
        digits = []
        for Jv, members in groups:
            if Jv != 0.0:
                digits.append(("g", 2 * len(members) + 1, (Jv, members)))
        for i in extras:
            digits.append(("e", 3, i))
        digits.append(("s", 2, None))
        size = int(np.prod([d[1] for d in digits]))
        layouts.append(digits)
        offsets.append(total)
        total += size
    return layouts, offsets, total


def ising_dE_values(sys_: System) -> np.ndarray:
    """dE of every Ising table entry, float32, evaluated in the reference's fp32 operation order: groups in
    order, v = v + fp32(J) * g, extras v = v + fp32(E) * t, field v = v + fp32(h), dE = -((t0 - s0) * v)
    with t0 = -s0 (SCK/generator.py:226-296)."""
    layouts, offsets, total = ising_layout(sys_)
    out = np.zeros(total, np.float32)
    f32 = np.float32
    for a in range(sys_.b):
        digits = layouts[a]
        sizes = [d[1] for d in digits]
        n_ent = int(np.prod(sizes))
        idx = np.arange(n_ent)
        vals = []
        for sz in reversed(sizes):               # last digit fastest
            vals.append(idx % sz)
            idx = idx // sz
        vals = vals[::-1]
        groups, extras = _groups(sys_, a)
        v = np.zeros(n_ent, np.float32)
        di = 0
        for Jv, members in groups:
            if Jv != 0.0:
                g = (vals[di] - len(members)).astype(np.float32)
                di += 1
                v = (v + f32(Jv) * g).astype(np.float32)
        ext_vals = []
        for i in extras:
            ext_vals.append((vals[di] - 1).astype(np.float32))
            di += 1
        for j, i in enumerate(extras):
            E = np.asarray(sys_.extra[i], np.float64).reshape(1, 1)
            if E[0, 0] != 0.0:
                v = (v + f32(E[0, 0]) * ext_vals[j]).astype(np.float32)
        if sys_.h is not None and sys_.h[0] != 0.0:
            v = (v + f32(sys_.h[0])).astype(np.float32)
        s0 = np.where(vals[di] > 0, f32(1.0), f32(-1.0)).astype(np.float32)
        t0 = (-s0).astype(np.float32)
        dE = (-((t0 - s0).astype(np.float32) * v).astype(np.float32)).astype(np.float32)
        out[offsets[a]:offsets[a] + n_ent] = dE
    return out


def ising_table(sys_: System, beta: np.ndarray) -> np.ndarray:
    """Per-replica acceptance table [Bp, entries], float32: Metropolis p = exp((-beta) dE), Barker
    q = 1 + exp(beta dE), with numpy's float32 exp, as the reference's CPU chain evaluates
    `u0 < exp(-beta dE)` and `u0 * (1 + exp(beta dE)) < 1` (SCK/generator.py:303-309)."""
    dE = ising_dE_values(sys_)
    beta = np.asarray(beta, np.float32).reshape(-1, 1)
    with np.errstate(over="ignore"):
        if sys_.acceptance == "barker":
            x = (beta * dE[None, :]).astype(np.float32)
            return (np.float32(1.0) + np.exp(x)).astype(np.float32)
        x = ((-beta) * dE[None, :]).astype(np.float32)
        return np.exp(x).astype(np.float32)


# --------------------------------------------------------------------------------------------- source
HEADER = r"""
// hipRTC provides the HIP device runtime itself, so only an offline hipcc build names the runtime header: a
// run-time compile then does not depend on the HIP install's header path (a profiler's environment can hide it)
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

typedef struct sck_args {
    void *spins; void *dst; const int8_t *mask; const float *k1; const float *beta; const void *table;
    const int64_t *state; const float *width; const float *site_ku; const float *site_axis;
    double *obs; double *obs_scratch; int64_t sweep_offset; int32_t nrep; int32_t nsweep;
    const void *aux; const int32_t *tiles; int32_t ntile; int32_t rep_base;
    double *track; double *track_parts;
} sck_args;

#define DEV static __device__ __forceinline__

// four floats as a native vector (registers, one 16-byte access)
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Philox-4x32-10 (Salmon et al., SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments
// 0x9E3779B9 / 0xBB67AE85; one 32x32->64 multiply gives each product's high and low words.
DEV void philox10(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        const uint64_t p2 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c1, k0, 0x96);   // 3-input xor
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p2 >> 32), c3, k1, 0x96);
        c1 = (uint32_t)p0;
        c3 = (uint32_t)p2;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// philox10 for the counter (c0, c1, 0, 0) and key (k0, 0), the form every sweep draw uses (c1 = 0 unless the
// counter is 64-bit). The first Philox round has one product (the other multiplies c2 = 0) and leaves
// c0 = c1 ^ k0, c1 = 0, so the second round's second product depends on c1 and the key alone and is uniform
// across a wave when they are; the third to tenth rounds are the general ones.
DEV void philox10_z(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0) {
    const uint64_t q = (uint64_t)0xD2511F53u * (uint64_t)c0;
    const uint64_t p0 = (uint64_t)0xCD9E8D57u * (uint64_t)(uint32_t)(q >> 32);
    const uint64_t pk = (uint64_t)0xD2511F53u * (uint64_t)(c1 ^ k0);
    c0 = (uint32_t)(p0 >> 32) ^ (k0 + 0x9E3779B9u);
    c1 = (uint32_t)p0;
    c2 = (uint32_t)q ^ ((uint32_t)(pk >> 32) ^ 0xBB67AE85u);
    c3 = (uint32_t)pk;
    uint32_t k = k0 + 2u * 0x9E3779B9u, l = 2u * 0xBB67AE85u;
#pragma unroll
    for (int r = 2; r < 10; ++r) {
        const uint64_t a = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        const uint64_t b = (uint64_t)0xD2511F53u * (uint64_t)c0;
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(a >> 32), c1, k, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(b >> 32), c3, l, 0x96);
        c1 = (uint32_t)a;
        c3 = (uint32_t)b;
        c0 = n0;
        c2 = n2;
        k += 0x9E3779B9u;
        l += 0xBB67AE85u;
    }
}

// uniform in [0, 1): fp32(x < 0 ? ~x : x) * fp32(4.6566127342e-10), x the output read as int32
DEV float u01(uint32_t x) {
    int32_t v = (int32_t)x;
    v = v < 0 ? ~v : v;
    return (float)v * 4.6566127342e-10f;
}

// per-sweep Philox key (base_seed * 1000003 + sweep * 7919 + 12345) mod 2147483629
DEV uint32_t sweep_key(const sck_args &A) {
    const int64_t base = A.state[0];
    const int64_t sweep = A.state[1] + A.sweep_offset;
    return (uint32_t)((base * 1000003LL + sweep * 7919LL + 12345LL) % 2147483629LL);
}

// sin and cos of a bounded argument |x| <= 16: Cody-Waite reduction by pi/2 with a two-part constant, then the
// Cephes single-precision minimax polynomials on [-pi/4, pi/4] (S. Moshier, Cephes Math Library, sinf.c).
DEV void sincos_bounded(float x, float *s_out, float *c_out) {
    const float k = __builtin_rintf(x * 0.63661977236758134f);
    float y = __builtin_fmaf(-k, 1.5707963705062866f, x);
    y = __builtin_fmaf(-k, -4.3711388286737929e-08f, y);
    const float y2 = y * y;
    float sp = __builtin_fmaf(y2, -1.9515295891e-4f, 8.3321608736e-3f);
    sp = __builtin_fmaf(y2, sp, -1.6666654611e-1f);
    const float s = __builtin_fmaf(y * y2, sp, y);
    float cp = __builtin_fmaf(y2, 2.443315711809948e-5f, -1.388731625493765e-3f);
    cp = __builtin_fmaf(y2, cp, 4.166664568298827e-2f);
    const float c = __builtin_fmaf(y2 * y2, cp, __builtin_fmaf(-0.5f, y2, 1.0f));
    const uint32_t q = (uint32_t)(int)k;
    const float ss = (q & 1u) ? c : s, cc = (q & 1u) ? s : c;
    // the quadrant's signs as xors of the sign bit (bit 1 of q, and of q + 1, shifted to bit 31): the same bits
    // as negating, in two instructions each instead of a test and a select
    *s_out = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, ss) ^ ((q << 30) & 0x80000000u));
    *c_out = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, cc) ^ (((q + 1u) << 30) & 0x80000000u));
}

// e^x through the hardware base-2 exponential
DEV float exp_fast(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

DEV int wrap_add(int v, int d, int L) {
    v += d;
    v = v >= L ? v - L : v;
    return v < 0 ? v + L : v;
}
"""


class _Gen:
    def __init__(self, sys_: System, mode: int, B: int, proposal: str, learned_meta=None, replica_major=False,
                 rm_tile=None, rm_chunks=None, rm_sites=None, resident_threads=None, rm_block=None, tiled=False,
                 tile_dims=None):
        self.s = sys_
        self.want_tiled = bool(tiled)
        self.tile_dims = None if tile_dims is None else tuple(int(v) for v in tile_dims)
        if self.tile_dims is not None and len(self.tile_dims) == 3:
            # (x, y, threads): a workgroup of 512 or 1024 threads asked for by name may take a tile of up to
            # 80 KB, so that two of them share a CU's 160 KB of LDS at four or eight waves per SIMD
            th = self.tile_dims[2]
            if th not in (256, 512, 1024):
                raise ValueError("tiled sweep workgroups have 256, 512 or 1024 threads")
            self.TILE_THREADS, self.tile_dims = th, self.tile_dims[:2]
            if th >= 512:
                self.TILE_LDS_MAX = 80 * 1024
        elif self.tile_dims is not None and len(self.tile_dims) != 2:
            raise ValueError("tile_dims is (x, y) or (x, y, threads)")
        self.mode = mode
        self.B = int(B)
        self.proposal = proposal
        self.lm = learned_meta
        self.n, self.b = sys_.n, sys_.b
        self.L1, self.L2, self.L3 = sys_.L1, sys_.L2, sys_.L3
        self.D = sys_.crystal.D
        self.per = list(sys_.periodic) + [False] * (3 - self.D)
        self.PL = self.L3 * self.L2 * self.L1
        self.NS = self.b * self.PL
        self.ROWS = self.L3 * self.L2
        self.W = -(-self.B // WORD) if mode in (MODE_PACKED, MODE_BITWISE, MODE_BITWISE_INDEP) else self.B
        self.Bp = self.W * WORD if mode in (MODE_PACKED, MODE_BITWISE, MODE_BITWISE_INDEP) else self.B
        self.has_mask = sys_.mask is not None
        self.classes = sys_.classes()
        self.P = x_period(sys_.colouring)
        self.BX = -(-self.L1 // self.P)
        self.wide = self.Bp * self.NS >= 2 ** 31    # 64-bit Philox counter arithmetic needed
        # replica-major storage [b][L3][L2][L1][n][B] (lanes over replicas) or the reference layout
        self.replica_major = bool(replica_major) and mode in (MODE_INT8, MODE_F32)
        self.SS = self.B if self.replica_major else 1
        self.rm_tile = self.default_rm_tile() if rm_tile is None else tuple(int(v) for v in rm_tile)
        self.rm_R = None if rm_chunks is None else max(1, int(rm_chunks))
        # class sites per workgroup: 1 (the 4 waves share a site, 256 R replicas) or 4 (a site per wave, 64 R)
        self.rm_SP = 1 if rm_sites is None else int(rm_sites)
        if resident_threads is not None:
            rt = int(resident_threads)
            if rt < 64 or rt > 1024 or rt % 64:
                raise ValueError("resident_threads must be a multiple of 64 in [64, 1024]")
            self.RESIDENT_THREADS = rt
        else:
            # one workgroup per replica: as many threads as a class has sites, 256 to 1024, so that a batch of
            # a few hundred replicas still fills the SIMDs (C1: 1024 threads, 8 waves per SIMD, 1.5x faster
            # than 256)
            self.RESIDENT_THREADS = min(1024, max(256, 64 * -(-(self.ROWS * self.BX) // 64)))
        if self.rm_SP not in (1, BLOCK // 64):
            raise ValueError(f"replica-major sites per workgroup must be 1 or {BLOCK // 64}")
        esz = {MODE_INT8: 1, MODE_F64: 8}.get(mode, 4)
        self.rm_block = self.B
        if self.replica_major:
            if rm_block is not None and int(rm_block) not in (0, self.B):
                rb = int(rm_block)
                if rb < 64 or rb % 64 or self.B % rb:
                    raise ValueError("a replica block must be a multiple of 64 that divides the batch")
                self.rm_block = rb
            elif rm_block is None:
                # a lattice larger than the cache runs in replica blocks whose share stays cache-resident
                per = self.n * self.NS * esz
                if per * self.B > RM_BLOCK_BYTES:
                    fits = [c for c in range(64, self.B, 64) if self.B % c == 0 and c * per <= RM_BLOCK_BYTES]
                    if fits:
                        self.rm_block = max(fits)
        # one replica's spins addressed by a 32-bit byte offset from its base pointer
        self.u32 = mode in (MODE_INT8, MODE_F32, MODE_F64) and self.n * self.NS * esz < 2 ** 31
        self.rt = False               # box dimensions are compile-time constants (runtime: batch kernels)
        self.lines: list[str] = []

    # -------------------------------------------------------------------------------------- helpers
    def word(self, base: str, site: str, T: str = "uint32_t", const: bool = True) -> str:
        """Site `site` of one packed word column whose base pointer `base` is wave-uniform in the class
        kernels. The 32-bit byte offset lets the load or store take the scalar-base form (no per-lane 64-bit
        address arithmetic); it is exact while a column's 4 NS bytes stay below 2^32."""
        if 4 * self.NS >= 2 ** 32:
            return f"{base}[{site}]"
        c = "const " if const else ""
        return f"(*({c}{T} *)(({c}char *)({base}) + (uint32_t)({site}) * 4u))"

    def at(self, ptr: str, comp: int, site: str) -> str:
        """Component `comp` of site `site` of the replica whose offset is `rbase`. Replica-major storage is
        [site][n][B]: a wave-uniform byte address of the site (scalar) plus the lane's 32-bit offset
        (replica and component), so each load takes a scalar base and a vector offset."""
        if self.replica_major:
            return f"rm_load({ptr}, (int64_t)({site}), (uint32_t)rbase, {comp})"
        if self.u32:
            # the replica's base pointer (wave-uniform wherever the replica is) plus a 32-bit byte offset
            esz = {MODE_INT8: 1, MODE_F64: 8}.get(self.mode, 4)
            return (f"(*(decltype({ptr}))((const char *)({ptr} + rbase) + "
                    f"(uint32_t)({self._mul(comp, self.NS)} + {site}) * {esz}u))")
        return f"{ptr}[rbase + (int64_t)({self._mul(comp, self.NS)} + {site})]"

    def store(self, ptr: str, comp: int, site: str, value: str) -> str:
        """Statement writing `value` to component `comp` of site `site` of the replica at `rbase`."""
        if self.replica_major:
            return f"rm_store({ptr}, (int64_t)({site}), (uint32_t)rbase, {comp}, {value});"
        return f"{self.at(ptr, comp, site)} = {value};"

    def emit_rm_access(self) -> None:
        """Replica-major loads and stores as buffer instructions: one wave-uniform buffer resource per site
        (scalar registers), the component's offset as the scalar offset and the lane's replica as the only
        vector offset, so an access costs no vector address arithmetic."""
        w = self.w
        T = self.elt()
        esz = 1 if self.mode == MODE_INT8 else 4
        w(f"DEV __amdgpu_buffer_rsrc_t rm_site(const {T} *base, int64_t site) {{")
        w(f"    return __builtin_amdgcn_make_buffer_rsrc((void *)((const char *)base + site * {self.n * self.B * esz}), "
          f"0, 0x7fffffff, 0x00020000);")
        w("}")
        w(f"DEV {T} rm_load(const {T} *base, int64_t site, uint32_t rep, int comp) {{")
        if esz == 4:
            w(f"    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rm_site(base, site), rep * 4u, "
              f"comp * {self.B * 4}, 0));")
        else:
            w(f"    return (int8_t)__builtin_amdgcn_raw_buffer_load_b8(rm_site(base, site), rep, comp * {self.B}, 0);")
        w("}")
        w(f"DEV void rm_store({T} *base, int64_t site, uint32_t rep, int comp, {T} v) {{")
        if esz == 4:
            w(f"    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rm_site(base, site), rep * 4u, "
              f"comp * {self.B * 4}, 0);")
        else:
            w(f"    __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v, rm_site(base, site), rep, comp * {self.B}, 0);")
        w("}")
        w()

    def rbase_expr(self, rep: str) -> str:
        """Offset of replica `rep`'s first spin: the address of component k of site q is
        rbase + (k * NS + q) * SS."""
        return f"(int64_t){rep}" if self.replica_major else f"(int64_t){rep} * {self._mul(self.n, self.NS)}"

    def w(self, s: str = "") -> None:
        self.lines.append(s)

    def elt(self) -> str:
        """The C type of one stored spin component (unpacked storage)."""
        return {MODE_INT8: "int8_t", MODE_F64: "double"}.get(self.mode, "float")

    def _mul(self, k: int, X) -> str:
        """k * X for a compile-time X (folded) or a runtime one (an expression of the batch dimensions)."""
        if isinstance(X, (int, np.integer)):
            return str(int(k) * int(X))
        return "0" if k == 0 else (X if k == 1 else f"{k} * {X}")

    def dims_param(self) -> str:
        """Extra parameter of the device helpers when the box is a runtime value."""
        return ", const bdims &D" if self.rt else ""

    def dims_arg(self) -> str:
        return ", D" if self.rt else ""

    def coord(self, var: str, d: int, L: int, periodic: bool):
        """(expression, validity or None) of a neighbour coordinate (SCK/generator.py:48-55)."""
        if d == 0:
            return var, None
        if periodic:
            if isinstance(L, (int, np.integer)) and L & (L - 1) == 0:
                # a power-of-two side wraps by masking (two's complement: exact for either sign of d)
                return f"(({var} + {d}) & {L - 1})", None
            if not isinstance(L, (int, np.integer)) or abs(d) < L:
                return f"wrap_add({var}, {d}, {L})", None
            return f"((({var} + {d}) % {L} + {L}) % {L})", None
        e = f"({var} + {d})"
        return e, f"({e} >= 0 && {e} < {L})"

    def nb_index(self, i: int):
        """Code computing q (neighbour site index) and ok (inside the box) of bond i from x, y, z."""
        bd = self.s.crystal.bonds[i]
        d = list(bd.delta) + [0] * (3 - self.D)
        ex, vx = self.coord("x", d[0], self.L1, self.per[0])
        ey, vy = self.coord("y", d[1], self.L2, self.per[1])
        ez, vz = self.coord("z", d[2], self.L3, self.per[2])
        valid = " && ".join(p for p in (vx, vy, vz) if p) or "true"
        return f"{self._mul(bd.alpha2, self.PL)} + (({ez}) * {self.L2} + ({ey})) * {self.L1} + ({ex})", valid

    def colour_expr(self, xv: str, a: int) -> str:
        """Colour of site (xv, y, z) of sublattice a (mixed radix of gamma(c, alpha))."""
        col = self.s.colouring
        if not col.m:
            return "0"
        expr = None
        for r, mr in enumerate(col.m):
            K = list(col.K[r]) + [0] * (3 - self.D)
            term = f"((({K[0]} * {xv} + {K[1]} * y + {K[2]} * z + {col.o[a][r]}) % {mr}))"
            expr = term if expr is None else f"(({expr}) * {mr} + {term})"
        return expr

    # ------------------------------------------------------------------------ Ising table decisions
    def decided_expr(self, p: str) -> str:
        """True where the table value alone decides the move (no draw needed): Metropolis p >= 1 (every
        u0 <= 1 - 2^-24 is below it) or p == 0 (u0 >= 0 never is); Barker q == 1 (u0 * 1 < 1) or q == inf."""
        if self.s.acceptance == "barker":
            return f"({p} == 1.0f || !({p} < __builtin_inff()))"
        return f"({p} >= 1.0f || !({p} > 0.0f))"

    def decided_value(self, p: str) -> str:
        return f"({p} == 1.0f ? 1 : 0)" if self.s.acceptance == "barker" else f"({p} >= 1.0f ? 1 : 0)"

    def accept_expr(self, p: str) -> str:
        return f"((u0 * {p}) < 1.0f)" if self.s.acceptance == "barker" else f"(u0 < {p})"

    # ------------------------------------------------------------------------------- site update
    def emit_update_plain(self, a: int) -> None:
        """__device__ update of one site of sublattice a for one replica (unpacked storage).
        Returns 1 when the move is accepted. `src` is read, `dst` written (in place: src == dst)."""
        s, n = self.s, self.n
        T = self.elt()
        w = self.w
        groups, extras = _groups(s, a)
        if n == 1:
            layouts, offsets, total = ising_layout(s)
            digits = layouts[a]
            # table lookup: p (Metropolis) or 1 + exp(beta dE) (Barker) of the site's neighbour configuration
            w(f"DEV float isp_{a}(const {T} *__restrict__ src, int64_t rbase, int own, int x, int y, int z, "
              f"const float *__restrict__ tb, int *s_out{self.dims_param()}) {{")
            w(f"    const int s0 = (int){self.at('src', 0, 'own')};")
            w("    *s_out = s0;")
            ng = sum(1 for Jv, _ in groups if Jv != 0.0)
            for k in range(ng):
                w(f"    int gsum{k} = 0;")
            for k in range(len(extras)):
                w(f"    int ext{k} = 0;")
            gi = 0
            for Jv, members in groups:
                w("    {")
                w("        int g = 0;")
                for i in members:
                    q, ok = self.nb_index(i)
                    if ok == "true":
                        w(f"        {{ const int q = {q}; const int t = (int){self.at('src', 0, 'q')}; g += t;"
                          + (f" ext{extras.index(i)} = t;" if i in extras else "") + " }")
                    else:
                        # outside the box: read the site itself (always in range) and use 0, so the loads
                        # issue without branches
                        w(f"        {{ const bool ok = {ok}; const int q = ok ? ({q}) : own; "
                          f"const int r = (int){self.at('src', 0, 'q')}; const int t = ok ? r : 0; g += t;"
                          + (f" ext{extras.index(i)} = t;" if i in extras else "") + " }")
                if Jv != 0.0:
                    w(f"        gsum{gi} = g + {len(members)};")
                    gi += 1
                w("    }")
            terms, gk, ek = [], 0, 0
            for kind, size, payload in digits:
                if kind == "g":
                    terms.append((f"gsum{gk}", size))
                    gk += 1
                elif kind == "e":
                    terms.append((f"(ext{ek} + 1)", size))
                    ek += 1
                else:
                    terms.append(("(s0 > 0 ? 1 : 0)", size))
            expr = "0"
            for t, size in terms:
                expr = f"({expr}) * {size} + {t}"
            w(f"    return tb[{offsets[a]} + {expr}];")
            w("}")
            w()
            w(f"DEV int upd_{a}(const sck_args &A, const {T} *__restrict__ src, {T} *__restrict__ dst, int64_t rbase, "
              f"int rep, int own, int x, int y, int z, uint32_t key{self.dims_param()}) {{")
            w("    int s0;")
            w(f"    const float p = isp_{a}(src, rbase, own, x, y, z, (const float *)A.table + (int64_t)rep * {total}, "
              f"&s0{self.dims_arg()});")
            w("    int acc;")
            w(f"    if ({self.decided_expr('p')}) acc = {self.decided_value('p')};")
            w("    else {")
            self.emit_draw("        ")
            w(f"        acc = {self.accept_expr('p')};")
            w("    }")
            w(f"    if (acc) {{ {self.store('dst', 0, 'own', f'({T})(-s0)')} }}")
            w("    return acc;")
            w("}")
            return
        w(f"DEV int upd_{a}(const sck_args &A, const {T} *__restrict__ src, {T} *__restrict__ dst, int64_t rbase, "
          f"int rep, int own, int x, int y, int z, uint32_t key{self.dims_param()}) {{")
        # vector models: fp32 local field in the reference's order
        for k in range(n):
            w(f"    float s{k} = (float){self.at('src', k, 'own')};")
        self.emit_local_field(a)
        self.emit_vector_tail(lambda k, v: self.store('dst', k, 'own', v))
        w("    return acc;")
        w("}")

    def emit_vector_tail(self, store, beta: str = "A.beta[rep]", width: str = "A.width[rep]") -> None:
        """Body of a vector-model update after the local field v<k> of own spin s<k>: draw, proposal t<k>,
        dE in the reference's fp32 order (SCK/generator.py:274-309), the anisotropy terms, and the Metropolis
        or Barker decision `acc`; store(k, value) gives the statement writing component k of the accepted
        trial spin."""
        s, n, w = self.s, self.n, self.w
        self.emit_draw("    ")
        self.emit_proposal(width, beta)
        dsum = f"(t0 - s0) * v0"
        for k in range(1, n):
            dsum = f"({dsum}) + (t{k} - s{k}) * v{k}"
        w(f"    float dE = -({dsum});")
        if s.cubic_k1 is not None:
            k2 = _hexf(s.cubic_k2)
            w("    { const float k1 = A.k1[own];")
            for p in ("t", "s"):
                w(f"      const float a_{p} = k1 * ((({p}0 * {p}1) * ({p}0 * {p}1) + ({p}0 * {p}2) * ({p}0 * {p}2)) "
                  f"+ ({p}1 * {p}2) * ({p}1 * {p}2)) + ({k2} * (({p}0 * {p}1) * {p}2)) * (({p}0 * {p}1) * {p}2);")
            w("      dE = dE + (a_t - a_s); }")
        if s.uniaxial is not None:
            Du, e = s.uniaxial
            w(f"    {{ const float se = {self._lassoc([f's{k} * {_hexf(e[k])}' for k in range(n)])};"
              f" const float te = {self._lassoc([f't{k} * {_hexf(e[k])}' for k in range(n)])};")
            w(f"      dE = dE + {_hexf(Du)} * (se * se - te * te); }}")
        if s.site_ku is not None:
            # only sites with a surface term read its axis: where ku == 0 the term adds a signed zero, which
            # leaves exp(-beta dE) (and Barker's 1 + exp(beta dE)) unchanged
            w(f"    {{ const float ku = A.site_ku[own]; if (ku != 0.0f) {{ const float *ax = A.site_axis + (int64_t)own * {n};")
            w(f"      const float sn = {self._lassoc([f's{k} * ax[{k}]' for k in range(n)])};"
              f" const float tn = {self._lassoc([f't{k} * ax[{k}]' for k in range(n)])};")
            w("      dE = dE + ku * (sn * sn - tn * tn); } }")
        w(f"    const float beta = {beta};")
        if s.acceptance == "barker":
            w("    const int acc = (u0 * (1.0f + exp_fast(beta * dE))) < 1.0f;")
        else:
            w("    const int acc = u0 < exp_fast((-beta) * dE);")
        w("    if (acc) {")
        for k in range(n):
            w(f"        {store(k, f't{k}')}")
        if self.mode == MODE_F64:
            # the reference stores where(acc, t, s) cast to the storage type, s being the spin as loaded in
            # float32 (SCK/generator.py:310-319): a rejected site keeps its value rounded to float32
            w("    } else {")
            for k in range(n):
                w(f"        {store(k, f's{k}')}")
        w("    }")

    def emit_local_field(self, a: int, ind: str = "    ", nb=None) -> None:
        """v0..v{n-1}: the local field of site (x, y, z) of sublattice a read from `src` at replica offset
        `rbase`, in the reference's fp32 order: J groups in first-appearance order (v = v + J g), then the
        extra matrices of the extra bonds in bond order, then the field (SCK/generator.py:222-277).
        nb(i, k), when given, is the expression of component k of bond i's neighbour (always present: a tile in
        LDS holds zeros outside an open box).

        A sum starts from its first term rather than from 0.0f + that term (which the compiler may not fold): the
        two differ only in the sign of a zero sum, which changes neither dE's test (exp(-beta * (+-0)) = 1) nor
        any spin."""
        s, n, w = self.s, self.n, self.w
        groups, extras = _groups(s, a)
        for k in range(n):
            w(f"{ind}float v{k} = 0.0f;")
        for j in range(len(extras)):
            for k in range(n):
                w(f"{ind}float n{j}_{k} = 0.0f;")
        v_set = False
        for Jv, members in groups:
            w(f"{ind}{{")
            for k in range(n):
                w(f"{ind}    float g{k};")
            for mi, i in enumerate(members):
                add = (lambda k: f"g{k} = t{k};") if mi == 0 else (lambda k: f"g{k} = g{k} + t{k};")
                if nb is not None:
                    w(f"{ind}    {{")
                    for k in range(n):
                        w(f"{ind}      const float t{k} = {nb(i, k)}; {add(k)}")
                        if i in extras:
                            w(f"{ind}      n{extras.index(i)}_{k} = t{k};")
                    w(f"{ind}    }}")
                    continue
                q, ok = self.nb_index(i)
                if ok == "true":
                    w(f"{ind}    {{ const int q = {q};")
                else:
                    # outside the box: read the site itself (always in range) and use 0, so the neighbour loads
                    # issue together without branches
                    w(f"{ind}    {{ const bool ok = {ok}; const int q = ok ? ({q}) : own;")
                for k in range(n):
                    if ok == "true":
                        w(f"{ind}      const float t{k} = (float){self.at('src', k, 'q')}; {add(k)}")
                    else:
                        w(f"{ind}      const float r{k} = (float){self.at('src', k, 'q')};")
                        w(f"{ind}      const float t{k} = ok ? r{k} : 0.0f; {add(k)}")
                    if i in extras:
                        w(f"{ind}      n{extras.index(i)}_{k} = t{k};")
                w(f"{ind}    }}")
            if Jv != 0.0:
                for k in range(n):
                    if v_set:
                        w(f"{ind}    v{k} = v{k} + {_hexf(Jv)} * g{k};")
                    else:
                        w(f"{ind}    v{k} = {_hexf(Jv)} * g{k};")
                v_set = True
            w(f"{ind}}}")
        for j, i in enumerate(extras):
            E = np.asarray(s.extra[i], np.float64)
            for k in range(n):
                terms = [f"{_hexf(E[k, l])} * n{j}_{l}" for l in range(n) if E[k, l] != 0.0]
                if terms:
                    w(f"{ind}v{k} = v{k} + ({self._lassoc(terms)});")
        if s.h is not None:
            for k in range(n):
                if s.h[k] != 0.0:
                    w(f"{ind}v{k} = v{k} + {_hexf(s.h[k])};")

    @staticmethod
    def _lassoc(terms):
        acc = terms[0]
        for t in terms[1:]:
            acc = f"({acc}) + {t}"
        return acc

    def emit_draw(self, ind: str) -> None:
        w = self.w
        if self.wide and not self.rt and self.NS & (self.NS - 1) == 0:
            # NS = 2^k: rep * NS has k low zero bits and own < NS, so no carry reaches the high word, which is
            # rep >> (32 - k): uniform wherever the replica is (the tiled kernel), so Philox's first rounds run
            # partly on the scalar unit
            k = self.NS.bit_length() - 1
            w(f"{ind}uint32_t c0 = (uint32_t)rep * {self.NS}u + (uint32_t)own, c1 = (uint32_t)rep >> {32 - k}, "
              f"c2 = 0u, c3 = 0u;")
        elif self.wide:
            ns = f"{self.NS}ull" if not self.rt else f"(uint64_t){self.NS}"
            w(f"{ind}const uint64_t ctr = (uint64_t)rep * {ns} + (uint64_t)own;")
            w(f"{ind}uint32_t c0 = (uint32_t)ctr, c1 = (uint32_t)(ctr >> 32), c2 = 0u, c3 = 0u;")
        else:
            ns = f"{self.NS}u" if not self.rt else f"(uint32_t){self.NS}"
            w(f"{ind}uint32_t c0 = (uint32_t)rep * {ns} + (uint32_t)own, c1 = 0u, c2 = 0u, c3 = 0u;")
        w(f"{ind}philox10_z(c0, c1, c2, c3, key);")
        w(f"{ind}const float u0 = u01(c0);")
        if self.n > 1:
            w(f"{ind}const float u1 = u01(c1);")
            w(f"{ind}const float u2 = u01(c2);")

    def emit_proposal(self, width: str = "A.width[rep]", beta: str = "A.beta[rep]") -> None:
        """Trial spin t from (u1, u2): uniform redraw (SCK/generator.py:281-289) or a cone of half-angle w
        (`width`: the fixed cone's half-angle of this replica, `beta`: its inverse temperature)."""
        w, n = self.w, self.n
        TWO_PI = "6.283185307179586f"
        if self.proposal == "uniform":
            if n == 2:
                w(f"    const float phi = {TWO_PI} * u1;")
                w("    float t0, t1; sincos_bounded(phi, &t1, &t0);")
            else:
                w("    const float cz = 2.0f * u1 - 1.0f;")
                w(f"    const float phi = {TWO_PI} * u2;")
                w("    const float rr = sqrtf(fmaxf(1.0f - cz * cz, 0.0f));")
                w("    float sp, cp; sincos_bounded(phi, &sp, &cp);")
                w("    const float t0 = rr * cp, t1 = rr * sp, t2 = cz;")
            return
        # cone proposal
        if self.proposal == "cone":
            w(f"    const float wd = {width};")
        else:
            # learned width: the frozen network tabulated on quantised features of the site's environment
            lm = self.lm
            vv = self._lassoc([f"v{k} * v{k}" for k in range(n)])
            w(f"    const float hv = sqrtf({vv});")
            w(f"    const float bt = {beta};")
            kexpr = _hexf(abs(self.s.uniaxial[0]) if self.s.uniaxial is not None else 0.0)
            if self.s.site_ku is not None:
                kexpr = f"({kexpr} + fabsf(A.site_ku[own]))"
            if self.s.cubic_k1 is not None:
                kexpr = f"({kexpr} + fabsf(A.k1[own]))"
            w(f"    const float f0 = hv * bt;")
            # a feature that is zero for every finite field is emitted as the constant (strict fp32 keeps the
            # compiler from folding 0 * v, and f2 would cost a division): a model without anisotropy has f1 = 0,
            # and one whose spins have no component along the axis (XY with the z axis) f2 = 0
            zero1 = self.s.uniaxial is None and self.s.site_ku is None and self.s.cubic_k1 is None
            zero2 = all(float(lm['axis'][k]) == 0.0 for k in range(n))
            w("    const float f1 = 0.0f;" if zero1 else f"    const float f1 = {kexpr} * bt;")
            ve = self._lassoc([f"v{k} * {_hexf(lm['axis'][k])}" for k in range(n)])
            w("    const float f2 = 0.0f;" if zero2 else f"    const float f2 = hv > 0.0f ? fabsf({ve}) / hv : 0.0f;")
            for d in range(3):
                if (d == 1 and zero1) or (d == 2 and zero2):
                    w(f"    const int i{d} = 0;")
                    continue
                w(f"    int i{d} = (int)(f{d} * {_hexf(lm['s'][d])}); i{d} = i{d} > {lm['n'][d] - 1} ? {lm['n'][d] - 1} : (i{d} < 0 ? 0 : i{d});")
            w(f"    const float wd = A.width[(i0 * {lm['n'][1]} + i1) * {lm['n'][2]} + i2];")
        if n == 2:
            w("    const float ph = atan2f(s1, s0) + wd * (2.0f * u1 - 1.0f);")
            w("    float t0, t1; sincos_bounded(ph, &t1, &t0);")
        else:
            w("    float sw, cw; sincos_bounded(wd, &sw, &cw);")
            w("    const float cc = 1.0f - u1 * (1.0f - cw);")
            w("    const float st = sqrtf(fmaxf(1.0f - cc * cc, 0.0f));")
            w(f"    const float phi = {TWO_PI} * u2;")
            w("    const bool nz = fabsf(s2) > 0.9f;")
            w("    const float ax0 = nz ? 1.0f : 0.0f, ax1 = 0.0f, ax2 = nz ? 0.0f : 1.0f;")
            w("    float e10 = ax1 * s2 - ax2 * s1, e11 = ax2 * s0 - ax0 * s2, e12 = ax0 * s1 - ax1 * s0;")
            w("    const float ne = sqrtf(e10 * e10 + e11 * e11 + e12 * e12);")
            w("    e10 = e10 / ne; e11 = e11 / ne; e12 = e12 / ne;")
            w("    const float e20 = s1 * e12 - s2 * e11, e21 = s2 * e10 - s0 * e12, e22 = s0 * e11 - s1 * e10;")
            w("    float sp, cp; sincos_bounded(phi, &sp, &cp);")
            w("    float t0 = cc * s0 + st * (cp * e10 + sp * e20);")
            w("    float t1 = cc * s1 + st * (cp * e11 + sp * e21);")
            w("    float t2 = cc * s2 + st * (cp * e12 + sp * e22);")
            w("    const float nt = sqrtf(t0 * t0 + t1 * t1 + t2 * t2);")
            w("    t0 = t0 / nt; t1 = t1 / nt; t2 = t2 / nt;")

    # ------------------------------------------------------------------------------- plain kernels
    def emit_plain(self) -> None:
        w = self.w
        T = self.elt()
        if self.replica_major:
            self.emit_rm_access()
        for a in range(self.b):
            self.emit_update_plain(a)
            w()
        # class kernels
        for k, (c, a) in enumerate(self.classes):
            if self.replica_major:
                self.emit_rm_class(k, c, a, T)
                continue
            w(f'extern "C" __global__ void __launch_bounds__({BLOCK}) sck_class_{k}(sck_args A) {{')
            w("    const int t = blockIdx.x * blockDim.x + threadIdx.x;")
            w(f"    const int row = t / {self.BX};")
            w(f"    if (row >= {self.ROWS}) return;")
            w(f"    const int j = t - row * {self.BX};")
            w(f"    const int z = row / {self.L2}, y = row - (row / {self.L2}) * {self.L2};")
            self.emit_residue(c, a)
            w(f"    const int x = x0 + {self.P} * j;")
            w(f"    if (x >= {self.L1}) return;")
            w(f"    const int own = {a * self.PL} + row * {self.L1} + x;")
            if self.has_mask:
                w("    if (!A.mask[own]) return;")
            w("    const int rep = blockIdx.y + blockIdx.z * gridDim.y;")
            w("    if (rep >= A.nrep) return;")
            w("    const uint32_t key = sweep_key(A);")
            w(f"    {T} *S = ({T} *)A.spins;")
            w(f"    upd_{a}(A, S, S, (int64_t)rep * {self.n * self.NS}, rep, own, x, y, z, key);")
            w("}")
            w()
        # sequential chain: one thread per replica
        w(f'extern "C" __global__ void __launch_bounds__(64) sck_seq(sck_args A) {{')
        w("    const int rep = blockIdx.x * blockDim.x + threadIdx.x;")
        w("    if (rep >= A.nrep) return;")
        w("    const uint32_t key = sweep_key(A);")
        w(f"    {T} *S = ({T} *)A.spins;")
        w(f"    const int64_t rbase = {self.rbase_expr('rep')};")
        for c, a in self.classes:
            w(f"    for (int z = 0; z < {self.L3}; ++z) for (int y = 0; y < {self.L2}; ++y) for (int x = 0; x < {self.L1}; ++x) {{")
            w(f"        if ({self.colour_expr('x', a)} != {c}) continue;")
            w(f"        const int own = {a * self.PL} + (z * {self.L2} + y) * {self.L1} + x;")
            if self.has_mask:
                w("        if (!A.mask[own]) continue;")
            w(f"        upd_{a}(A, S, S, rbase, rep, own, x, y, z, key);")
            w("    }")
        w("}")
        w()
        if self.n > 1:
            self.emit_sample_kernel(T)
        # synchronous control: every site of sublattice a from the old state
        for a in range(self.b):
            w(f'extern "C" __global__ void __launch_bounds__({BLOCK}) sck_sync_{a}(sck_args A) {{')
            if self.replica_major:
                w("    const int rep = blockIdx.x * blockDim.x + threadIdx.x;")
                w("    const int t = blockIdx.y + blockIdx.z * gridDim.y;")
                w(f"    if (t >= {self.PL} || rep >= {self.B}) return;")
                w(f"    const int x = t % {self.L1}, row = t / {self.L1};")
                w(f"    const int z = row / {self.L2}, y = row % {self.L2};")
                w(f"    const int own = {a * self.PL} + t;")
                if self.has_mask:
                    w("    if (!A.mask[own]) return;")
                w("    const uint32_t key = sweep_key(A);")
                w(f"    upd_{a}(A, (const {T} *)A.spins, ({T} *)A.dst, (int64_t)rep, rep, own, x, y, z, key);")
                w("}")
                w()
                continue
            w("    const int t = blockIdx.x * blockDim.x + threadIdx.x;")
            w(f"    if (t >= {self.PL}) return;")
            w(f"    const int x = t % {self.L1}, row = t / {self.L1};")
            w(f"    const int z = row / {self.L2}, y = row % {self.L2};")
            w(f"    const int own = {a * self.PL} + t;")
            if self.has_mask:
                w("    if (!A.mask[own]) return;")
            w("    const int rep = blockIdx.y + blockIdx.z * gridDim.y;")
            w("    if (rep >= A.nrep) return;")
            w("    const uint32_t key = sweep_key(A);")
            w(f"    const int64_t rbase = {self.rbase_expr('rep')};")
            w(f"    upd_{a}(A, (const {T} *)A.spins, ({T} *)A.dst, rbase, rep, own, x, y, z, key);")
            w("}")
            w()

    def default_rm_tile(self):
        """Class slots per tile (x slots, rows, planes) of the replica-major kernels: a tile's sites and most of
        their neighbours stay within the L2 window of the workgroups one XCD holds in flight."""
        if self.L3 > 1:
            return (min(self.BX, 4), min(self.L2, 4), min(self.L3, 4))
        return (min(self.BX, 32), min(self.L2, 16), 1)

    def rm_plan(self):
        """(groups, workgroup slots per group, items per XCD) of the replica-major class kernels; slots are
        padded to whole tiles, and a workgroup takes rm_SP consecutive slots of one group."""
        tx, ty, tz = self.rm_tile
        ntile = -(-self.BX // tx) * -(-self.L2 // ty) * -(-self.L3 // tz)
        S = -(-(ntile * tx * ty * tz) // self.rm_SP)
        G = -(-self.rm_block // (BLOCK * self.rm_chunks() // self.rm_SP))
        return G, S, -(-(G * S) // XCDS)

    def emit_rm_class(self, k: int, c: int, a: int, T: str) -> None:
        """Replica-major class kernel: the 64 lanes of a wave are 64 consecutive replicas of one class site, so
        the site, its colour residue and every neighbour index are wave-uniform (scalar) and each spin load is
        one coalesced 256-byte access.

        Work items are (replica group, class slot) pairs. Workgroups are dispatched round-robin over the 8 XCDs,
        so blockIdx.x is the XCD and (blockIdx.y, blockIdx.z) the position inside that XCD's contiguous share
        of the items; slots run through 3-D tiles (x slots fastest, then rows, then planes), so a site's
        neighbours are mostly visited by workgroups of the same XCD shortly before or after it and come from
        that XCD's L2."""
        w = self.w
        R = self.rm_chunks()
        G, S, chunk = self.rm_plan()
        tx, ty, tz = self.rm_tile
        ntx, nty = -(-self.BX // tx), -(-self.L2 // ty)
        tsz = tx * ty * tz
        # Heisenberg: at most RM_WAVES waves per SIMD. Left free to target 8, the compiler keeps the cone kernel
        # at 31 registers by waiting on each neighbour load before issuing the next (9 loads in flight instead
        # of 37); XY kernels keep all their loads in flight at 8 waves and stay uncapped.
        cap = f" __attribute__((amdgpu_waves_per_eu(1, {RM_WAVES})))" if self.n == 3 else ""
        w(f'extern "C" __global__ void __launch_bounds__({BLOCK}){cap} sck_class_{k}(sck_args A) {{')
        w(f"    const int64_t v = (int64_t)blockIdx.x * {chunk} + blockIdx.y + (int64_t)blockIdx.z * gridDim.y;")
        w(f"    if (blockIdx.y + (int64_t)blockIdx.z * gridDim.y >= {chunk} || v >= {G * S}LL) return;")
        if self.rm_SP == 1:
            w(f"    const int grp = (int)(v / {S}), t = (int)(v - (int64_t)grp * {S});")
            w(f"    const int rep0 = A.rep_base + grp * {BLOCK * R} + (threadIdx.x >> 6) * {64 * R} + (threadIdx.x & 63);")
        else:
            # a site per wave: the wave index made scalar so the site arithmetic stays in scalar registers
            w("    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);")
            w(f"    const int grp = (int)(v / {S});")
            w(f"    const int t = (int)(v - (int64_t)grp * {S}) * {self.rm_SP} + wave;")
            w(f"    const int rep0 = A.rep_base + grp * {64 * R} + (threadIdx.x & 63);")
        w(f"    const int tt = t / {tsz}, u = t - tt * {tsz};")
        w(f"    const int tj = tt % {ntx}, tr = tt / {ntx};")
        w(f"    const int uy = (u / {tx}) % {ty}, uz = u / {tx * ty};")
        w(f"    const int j = tj * {tx} + (u % {tx});")
        w(f"    const int y = (tr % {nty}) * {ty} + uy, z = (tr / {nty}) * {tz} + uz;")
        w(f"    if (j >= {self.BX} || y >= {self.L2} || z >= {self.L3}) return;")
        w(f"    const int row = z * {self.L2} + y;")
        self.emit_residue(c, a)
        w(f"    const int x = x0 + {self.P} * j;")
        w(f"    if (x >= {self.L1}) return;")
        w(f"    const int own = {a * self.PL} + row * {self.L1} + x;")
        if self.has_mask:
            w("    if (!A.mask[own]) return;")
        w("    const uint32_t key = sweep_key(A);")
        w(f"    {T} *S = ({T} *)A.spins;")
        # one wave, one site, R chunks of 64 replicas: the site's scalar address work is done once
        w(f"    for (int i = 0; i < {R}; ++i) {{")
        w("        const int rep = rep0 + 64 * i;")
        # the launch covers replicas [rep_base, rep_base + block) (the whole batch when unblocked)
        w(f"        if (rep - A.rep_base < {self.rm_block}) upd_{a}(A, S, S, (int64_t)rep, rep, own, x, y, z, key);")
        w("    }")
        w("}")
        w()

    def emit_sample_kernel(self, T: str) -> None:
        """sck_sample: training records for the learned proposal, one per (replica, site) pair of rs: the
        spin, the local field of its neighbours, beta and the site's anisotropy constants (12 floats, the
        layout of sck_mlp_sample)."""
        w = self.w
        w(f'extern "C" __global__ void __launch_bounds__(128) sck_sample(sck_args A, const int32_t *rs, int32_t nsample, '
          f'float *out) {{')
        w("    const int i = blockIdx.x * blockDim.x + threadIdx.x;")
        w("    if (i >= nsample) return;")
        w("    const int rep = rs[2 * i], own = rs[2 * i + 1];")
        w(f"    const int a = own / {self.PL}, rem = own - a * {self.PL};")
        w(f"    const int x = rem % {self.L1}, row = rem / {self.L1};")
        w(f"    const int z = row / {self.L2}, y = row % {self.L2};")
        w(f"    const {T} *src = (const {T} *)A.spins;")
        w(f"    const int64_t rbase = {self.rbase_expr('rep')};")
        w("    float *o = out + (int64_t)i * 12;")
        w("    for (int k = 0; k < 12; ++k) o[k] = 0.0f;")
        for k in range(self.n):
            w(f"    o[{k}] = (float){self.at('src', k, 'own')};")
        for a in range(self.b):
            w(f"    if (a == {a}) {{")
            self.emit_local_field(a, "        ")
            for k in range(self.n):
                w(f"        o[{3 + k}] = v{k};")
            w("    }")
        w("    o[6] = A.beta[rep];")
        if self.s.cubic_k1 is not None:
            w("    o[7] = A.k1[own];")
        if self.s.site_ku is not None:
            w("    o[8] = A.site_ku[own];")
            for k in range(self.n):
                w(f"    o[{9 + k}] = A.site_axis[(int64_t)own * {self.n} + {k}];")
        w("}")
        w()

    def rm_chunks(self) -> int:
        """64-replica chunks one wave of a replica-major class kernel runs for its site: more chunks share the
        site's scalar work, fewer keep the workgroups in flight on an XCD within its L2."""
        if self.rm_R is not None:
            return self.rm_R
        return max(1, min(4, self.rm_block // BLOCK))

    def emit_residue(self, c: int, a: int) -> None:
        """x0: the x residue in [0, P) of row (y, z) where sublattice a has colour c (P when none)."""
        w = self.w
        w(f"    int x0 = {self.P};")
        for q in range(self.P - 1, -1, -1):
            w(f"    if ({self.colour_expr(str(q), a)} == {c}) x0 = {q};")
        w(f"    if (x0 >= {self.P}) return;")

    # ------------------------------------------------------------------------------- tiled sweep kernel
    # 512 threads: three 43.5 KB XY tiles of 128 x 32 cells per CU run six waves per SIMD instead of three
    # (C5 2.21e11 -> 2.41e11 site updates/s against 256 threads, same tile); the default tile budget stays at
    # 64 KB, so the tile each system gets is the one 256 threads had
    TILE_THREADS = 512
    TILE_LDS_MAX = 64 * 1024

    def tiled_plan(self):
        """(TX, TY, HX, HY) of the tiled sweep kernel, or None where it does not apply: float32 vector spins on a
        2-D lattice in the reference layout, with a tile and its halo fitting the LDS budget. The halo is the
        number of classes times the bond reach along each axis: class k of K is updated on the tile grown by
        (K - 1 - k) reaches, so that the K-th class pass is exact on the tile itself."""
        if self.mode != MODE_F32 or self.n < 2 or self.D != 2 or self.replica_major or not self.classes:
            return None
        rx = max((abs(bd.delta[0]) for bd in self.s.crystal.bonds), default=0)
        ry = max((abs(bd.delta[1]) for bd in self.s.crystal.bonds), default=0)
        K = len(self.classes)
        HX, HY = K * rx, K * ry
        vec = self.L1 % 4 == 0
        if vec:
            HX = -(-HX // 4) * 4              # whole float4 columns of halo (the extra cells are loaded only)
        # tile origins at multiples of the colouring's periods: every tile then holds the same class pattern,
        # so each class pass enumerates its sites from compile-time row tables
        px = math.lcm(self.P, 4 if vec else 1)
        py = y_period(self.s.colouring)
        for dims in ([self.tile_dims] if self.tile_dims is not None else TILE_CANDIDATES):
            TX, TY = min(dims[0], self.L1), min(dims[1], self.L2)
            TX, TY = TX // px * px, TY // py * py
            if TX == 0 or TY == 0:
                continue
            if self.b * (TX + 2 * HX) * (TY + 2 * HY) * self.n * 4 > self.TILE_LDS_MAX:
                continue
            if TX + 2 * HX > self.L1 or TY + 2 * HY > self.L2:
                continue                      # the grown tile would wrap onto itself: small lattices run per class
            return TX, TY, HX, HY, rx, ry
        return None

    def emit_tiled(self, plan) -> None:
        """sck_tiled: whole coloured sweeps, one launch per sweep. A workgroup takes (replica, tile) items: it loads
        the tile and a halo of K bond reaches into LDS, runs the K class passes in sweep order, class k on the tile
        grown by K - 1 - k reaches, with a barrier between passes, and writes the tile back. Every draw depends only
        on (seed, sweep, replica, site), so a halo site updated here and in the neighbouring tile takes the same
        value in both, and each tile's sites end exactly as after the K class launches. The lattice is read about
        (1 + 2 K r / T)^2 times and written once per sweep, against K reads and one write for per-class launches.

        A tile's halo cells belong to other tiles, which may finish first, so the sweep reads A.spins and writes
        A.dst (the launcher alternates the two buffers).

        Workgroups are persistent: gridDim.x = 8 is the XCD, gridDim.y the workgroups resident per XCD (the
        launcher sets it from the occupancy). The items of XCD x are one contiguous range, tiles of a replica in
        raster order, which its workgroups stride through, so a tile's halo rows were read shortly before by the
        tile above it and come from that XCD's L2. Each workgroup loads its next tile into registers while it
        runs the class passes of the current one. Where the box and halo are multiples of 4 cells along x, tiles
        move as float4 rows."""
        w, s, n, b = self.w, self.s, self.n, self.b
        TX, TY, HX, HY, rx, ry = plan
        W, H = TX + 2 * HX, TY + 2 * HY
        HW = H * W
        NTX, NTY = -(-self.L1 // TX), -(-self.L2 // TY)
        NT = NTX * NTY
        RT = self.TILE_THREADS
        K = len(self.classes)
        col = s.colouring
        V = 4 if (self.L1 % 4 == 0 and HX % 4 == 0 and TX % 4 == 0) else 1
        NCH = W // V                        # V-wide chunks per halo row (W is a multiple of V)
        NLD = b * H * NCH                   # chunks of one tile and its halo
        ULD = -(-NLD // RT)                 # chunks per thread
        NCI = TX // V if V > 1 else TX
        NST = b * TY * NCI                  # chunks of the tile itself
        VT = "f32x4" if V == 4 else "float"
        NC = b * HW                         # LDS cells per component
        # XY tiles whose rows move as float4 chunks keep a cell's two components side by side in LDS, so each
        # neighbour read is one 8-byte access at an immediate offset; other tiles are component-planar
        IL = n == 2 and V == 4

        def lds(cell: str, k: int) -> str:
            return f"ld[2 * ({cell}) + {k}]" if IL else f"ld[{k * NC} + {cell}]"
        for a in range(b):
            offs = {}
            for i, bd in enumerate(s.crystal.bonds):
                if bd.alpha == a:
                    offs[i] = (bd.alpha2 - a) * HW + bd.delta[1] * W + bd.delta[0]
            # beta and the cone width arrive per item (loaded once, ahead of the next tile's prefetch)
            # the move's energy change and spin change are reported for the running totals of E and M
            w(f"DEV int tupd_{a}(const sck_args &A, float *__restrict__ ld, int li, int rep, int own, uint32_t key, "
              f"float beta_r, float width_r, float &dE_out, float *ds_out) {{")
            for k in range(n):
                w(f"    float s{k} = {lds('li', k)};")
            self.emit_local_field(a, nb=lambda i, k, offs=offs: lds(f"li + ({offs[i]})", k))
            self.emit_vector_tail(lambda k, v: f"{lds('li', k)} = {v};", beta="beta_r", width="width_r")
            w("    dE_out = dE;")
            for k in range(n):
                w(f"    ds_out[{k}] = t{k} - s{k};")
            w("    return acc;")
            w("}")
            w()
        items = self.B * NT
        chunk = -(-items // XCDS)
        self.tiled_chunk = chunk

        def ax(var_u, L, periodic, out):
            # wrapped coordinate `out` and validity of unwrapped `var_u`
            if periodic:
                if L & (L - 1) == 0:
                    return f"const int {out} = {var_u} & {L - 1};", "true"
                return f"const int {out} = {var_u} < 0 ? {var_u} + {L} : ({var_u} >= {L} ? {var_u} - {L} : {var_u});", "true"
            return f"const int {out} = {var_u};", f"({var_u} >= 0 && {var_u} < {L})"

        cx, vx = ax("gxu", self.L1, self.per[0], "gx")
        cy, vy = ax("gyu", self.L2, self.per[1], "gy")
        zero = "(f32x4)(0.0f)" if V == 4 else "0.0f"
        # the LDS of a tile bounds the workgroups per CU (RT / 256 waves of each per SIMD): telling the compiler so lets
        # it keep the prefetched tile in registers instead of trading them for an occupancy it cannot reach
        wpe = max(1, min(8, (160 * 1024) // (b * HW * n * 4) * (RT // 256)))
        w(f'extern "C" __global__ void __launch_bounds__({RT}) __attribute__((amdgpu_waves_per_eu(1, {wpe}))) '
          f'sck_tiled(sck_args A) {{')
        w(f"    __shared__ __attribute__((aligned(16))) float ld[{b * HW * n}];")
        NV = 1 + b * n
        w(f"    __shared__ double red[{RT // 64}][{NV}];")
        w("    const bool trk = A.track_parts != nullptr;      // running totals of E and M (uniform)")
        w(f"    const int64_t lo = (int64_t)blockIdx.x * {chunk};")
        w(f"    const int64_t hi = lo + {chunk} < {items}LL ? lo + {chunk} : {items}LL;")
        w("    const int64_t step = gridDim.y;")
        w("    int64_t g = lo + blockIdx.y;")
        w("    if (g >= hi) return;")
        w("    const uint32_t key = sweep_key(A);")
        # one named register (vector) per prefetched chunk and component, fully unrolled here: an indexed array
        # would be placed in scratch memory
        w(f"    {VT} {', '.join(f'p{u}_{k}' for u in range(ULD) for k in range(n))};")
        def fetch(item: str, ind: str) -> None:
            """The loads of item `item` into p<u>_<k> (issued, not waited for). Emitted in place: a lambda called
            twice is compiled by hipRTC as a call whose closure lives in scratch memory."""
            w(f"{ind}{{")
            w(f"{ind}    const int64_t gi = {item};")
            w(f"{ind}    const int rep = (int)(gi / {NT}), tile = (int)(gi - (int64_t)rep * {NT});")
            w(f"{ind}    const int tx0 = (tile % {NTX}) * {TX}, ty0 = (tile / {NTX}) * {TY};")
            w(f"{ind}    const float *S = (const float *)A.spins + (int64_t)rep * {n * self.NS};")
            for u in range(ULD):
                w(f"{ind}    {{")
                w(f"{ind}        const int i = threadIdx.x + {u * RT};")
                w(f"{ind}        const int a = i / {H * NCH}, r = i - a * {H * NCH};")
                w(f"{ind}        const int ly = r / {NCH}, c = r - ly * {NCH};")
                w(f"{ind}        const int gxu = tx0 - {HX} + c * {V}, gyu = ty0 - {HY} + ly;")
                w(f"{ind}        {cx} {cy}")
                w(f"{ind}        const bool ok = {'i < ' + str(NLD) + ' && ' if (u + 1) * RT > NLD else ''}{vx} && {vy};")
                # an always-valid address, so the loads issue without branches; cells outside the box (and the
                # padding chunks) are zeroed or skipped when written to LDS, so nothing waits for these loads here
                w(f"{ind}        const int64_t q = ok ? (int64_t)a * {self.PL} + (int64_t)gy * {self.L1} + gx : 0;")
                for k in range(n):
                    w(f"{ind}        p{u}_{k} = *(const {VT} *)(S + {self.NS * k} + q);")
                w(f"{ind}    }}")
            w(f"{ind}}}")

        fetch("g", "    ")
        w("    for (; g < hi; g += step) {")
        w(f"        const int rep = (int)(g / {NT}), tile = (int)(g - (int64_t)rep * {NT});")
        w(f"        const int tx0 = (tile % {NTX}) * {TX}, ty0 = (tile / {NTX}) * {TY};")
        # a replica's beta and cone width, made wave-uniform here: the wait for the load falls before the next
        # tile's prefetch is issued, and the class passes hold them in scalar registers
        w("        const float beta_r = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane("
          "__builtin_bit_cast(int, A.beta[rep])));")
        if self.proposal == "cone":
            w("        const float width_r = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane("
              "__builtin_bit_cast(int, A.width[rep])));")
        else:
            w("        const float width_r = 0.0f;")
        w(f"        double tE = 0.0{''.join(f', tm{a}_{k} = 0.0' for a in range(b) for k in range(n))};")
        w("        __syncthreads();                          // the previous item's write-back has read the LDS")
        conds = [v for v in (vx, vy) if v != "true"]
        for u in range(ULD):
            w("        {")
            w(f"            const int i = threadIdx.x + {u * RT};")
            guard = (u + 1) * RT > NLD
            if guard:
                w(f"            if (i < {NLD}) {{")
            w(f"            const int a = i / {H * NCH}, r = i - a * {H * NCH};")
            w(f"            const int ly = r / {NCH}, c = r - ly * {NCH};")
            w(f"            float *dl = ld + {'2 * (' if IL else '('}{a_expr('a', HW)} + ly * {W} + c * {V});")
            if conds:
                # outside an open box the LDS holds zeros (no spin there)
                w(f"            const int gxu = tx0 - {HX} + c * {V}, gyu = ty0 - {HY} + ly;")
                w(f"            const bool ok = {' && '.join(conds)};")
            vals = [f"(ok ? p{u}_{k} : {zero})" if conds else f"p{u}_{k}" for k in range(n)]
            if IL:
                # four cells' (x, y) pairs as two float4 stores
                w(f"            const f32x4 q0 = {vals[0]}, q1 = {vals[1]};")
                w("            *(f32x4 *)dl = (f32x4){q0.x, q1.x, q0.y, q1.y};")
                w("            *(f32x4 *)(dl + 4) = (f32x4){q0.z, q1.z, q0.w, q1.w};")
            else:
                for k in range(n):
                    w(f"            *({VT} *)(dl + {k * NC}) = {vals[k]};")
            if guard:
                w("            }")
            w("        }")
        w("        __syncthreads();")
        w("        if (g + step < hi)                         // in flight during the class passes")
        fetch("g + step", "        ")
        P, Py = self.P, y_period(col)
        self.tiled_tables = []
        passes = []                          # (k, a, comment, first lines of the pass, inner-site conditions)
        for k, (c, a) in enumerate(self.classes):
            e = K - 1 - k
            x_lo, y_lo = HX - e * rx, HY - e * ry
            wx, hy = TX + 2 * e * rx, TY + 2 * e * ry
            # rows of the pass that hold class sites, by row residue mod Py, and their first x residue: the tile
            # origin (tx0, ty0) is a multiple of (P, Py), so these depend on the local coordinates only
            xb = (x_lo - HX) % P
            res = []
            for r in range(Py):
                yy = (y_lo - HY + r) % Py
                q0 = next((q for q in range(P) if col.colour(((xb + q) % P, yy)[:self.D], a) == c), None)
                if q0 is not None:
                    res.append((r, q0))
            self.tiled_tables.append((x_lo, y_lo, wx, hy, res))
            if not res:
                continue
            nv = len(res)
            bxk = max(-(-(wx - q0) // P) for _, q0 in res)
            nblk = -(-hy // Py)
            nitem = nblk * nv * bxk
            rsel, qsel = f"{res[-1][0]}", f"{res[-1][1]}"
            for i in range(nv - 2, -1, -1):
                rsel = f"(ri == {i} ? {res[i][0]} : {rsel})"
                qsel = f"(ri == {i} ? {res[i][1]} : {qsel})"
            head = [f"// class {k}: colour {c}, sublattice {a}, on the tile grown by {e} reaches",
                    f"for (int t = threadIdx.x; t < {nitem}; t += {RT}) {{",
                    f"    const int u = t / {bxk}, j = t - u * {bxk};"]
            if nv == 1:
                head.append(f"    const int ly = {y_lo + res[0][0]} + {Py} * u, q0 = {res[0][1]};")
            else:
                head.append(f"    const int m = u / {nv}, ri = u - m * {nv};")
                head.append(f"    const int ly = {y_lo} + {Py} * m + {rsel}, q0 = {qsel};")
            head.append(f"    const int lx = {x_lo} + q0 + {P} * j;")
            # a bound is tested only where the enumeration can pass it (the largest ly and lx it produces)
            ly_max = y_lo + Py * (nblk - 1) + max(r for r, _ in res)
            lx_max = x_lo + max(q for _, q in res) + P * (bxk - 1)
            outside = ([f"ly >= {y_lo + hy}"] if ly_max >= y_lo + hy else []) + \
                ([f"lx >= {x_lo + wx}"] if lx_max >= x_lo + wx else [])
            if outside:
                head.append(f"    if ({' || '.join(outside)}) continue;")
            head.append(f"    const int gxu = tx0 - {HX} + lx, gyu = ty0 - {HY} + ly;")
            head.append(f"    {cx} {cy}")
            conds = [v for v in (vx, vy) if v != "true"]
            if conds:
                head.append(f"    if (!({' && '.join(conds)})) continue;")
            head.append(f"    const int own = {a * self.PL} + gy * {self.L1} + gx;")
            if self.has_mask:
                head.append("    if (!A.mask[own]) continue;")
            head.append(f"    float dEo, dso[{n}];")
            head.append(f"    const int acc = tupd_{a}(A, ld, {a * HW} + ly * {W} + lx, rep, own, key, beta_r, "
                        f"width_r, dEo, dso);")
            # the running totals count each site once: its move in the tile itself (not the halo copies)
            inner = [f"gxu < {self.L1}", f"gyu < {self.L2}"]
            if e > 0:
                inner = [f"lx >= {HX}", f"lx < {HX + TX}", f"ly >= {HY}", f"ly < {HY + TY}"] + inner
            passes.append((a, head, inner))
        # the passes are emitted twice, behind the uniform test of trk: with the totals' bookkeeping inside one
        # loop, the compiler turns it into float64 adds and selects that every site executes (21 of 124 vector
        # instructions per site at XY triangular) even where no totals are kept
        for track in (True, False):
            w("        if (trk) {" if track else "        } else {")
            for a, head, inner in passes:
                for line in head:
                    w("            " + line)
                if track:
                    w(f"                if (acc && {' && '.join(inner)}) {{")
                    w("                    tE += (double)dEo;")
                    for k2 in range(n):
                        w(f"                    tm{a}_{k2} += (double)dso[{k2}];")
                    w("                }")
                else:
                    w("                (void)acc;")
                w("            }")
                w("            __syncthreads();")
        w("        }")
        # write the tile back (its unwrapped cells inside the box: a partial tile's overhang is another tile's)
        vals = ["tE"] + [f"tm{a}_{k}" for a in range(b) for k in range(n)]
        w("        if (trk) {                                  // one partial of E and M per (replica, tile)")
        for j, v in enumerate(vals):
            w(f"            {{ double x = {v}; for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);"
              f" if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][{j}] = x; }}")
        w("            __syncthreads();")
        w(f"            if (threadIdx.x < {NV}) {{")
        w("                double x = 0.0;")
        w(f"                for (int q = 0; q < {RT // 64}; ++q) x += red[q][threadIdx.x];")
        w(f"                A.track_parts[((int64_t)tile * {self.B} + rep) * {NV} + threadIdx.x] = x;")
        w("            }")
        w("        }")
        w(f"        float *D = (float *)A.dst + (int64_t)rep * {n * self.NS};")
        w(f"        for (int i = threadIdx.x; i < {NST}; i += {RT}) {{")
        w(f"            const int a = i / {TY * NCI}, r = i - a * {TY * NCI};")
        w(f"            const int uy = r / {NCI}, c = r - uy * {NCI};")
        w(f"            const int gx = tx0 + c * {V}, gy = ty0 + uy;")
        w(f"            if (gx >= {self.L1} || gy >= {self.L2}) continue;")
        w(f"            const float *sl = ld + {'2 * (' if IL else '('}{a_expr('a', HW)} + (uy + {HY}) * {W} + c * {V} + {HX});")
        w(f"            const int64_t q = (int64_t)a * {self.PL} + (int64_t)gy * {self.L1} + gx;")
        # non-temporal: the written tile is not read again in this sweep, so its lines do not displace the halo
        # rows that the next tiles read from L2 (+2 % at XY 2048^2 x 2048, one box)
        if IL:
            w("            const f32x4 r0 = *(const f32x4 *)sl, r1 = *(const f32x4 *)(sl + 4);")
            w("            __builtin_nontemporal_store((f32x4){r0.x, r0.z, r1.x, r1.z}, (f32x4 *)(D + q));")
            w(f"            __builtin_nontemporal_store((f32x4){{r0.y, r0.w, r1.y, r1.w}}, (f32x4 *)(D + {self.NS} + q));")
        else:
            for k in range(n):
                w(f"            __builtin_nontemporal_store(*(const {VT} *)(sl + {k * NC}), ({VT} *)(D + {self.NS * k} + q));")
        w("        }")
        w("    }")
        w("}")
        w()
        # the running totals: one thread per (replica, value) adds that value's partials of every tile in tile
        # order (a fixed order: the same totals run to run)
        self.tiled_nt = NT
        w('extern "C" __global__ void __launch_bounds__(256) sck_track_fold(sck_args A) {')
        w("    const int i = blockIdx.x * blockDim.x + threadIdx.x;")
        w(f"    if (i >= {self.B * NV}) return;")
        w(f"    const int rep = i / {NV}, v = i - rep * {NV};")
        w("    double x = 0.0;")
        w(f"    for (int t = 0; t < {NT}; ++t) x += A.track_parts[((int64_t)t * {self.B} + rep) * {NV} + v];")
        w(f"    if (v == 0) A.track[rep] += x; else A.track[{self.B} + (int64_t)rep * {NV - 1} + v - 1] += x;")
        w("}")
        w()

    # ---------------------------------------------------------------------------- resident kernel
    RESIDENT_THREADS = 256           # overridden per module (see __init__)
    RESIDENT_LDS_MAX = 60 * 1024

    def resident_plan(self):
        """(chunk, LDS bytes) of the LDS-resident kernel, or None when one replica's lattice does not fit."""
        if self.mode not in (MODE_INT8, MODE_F32) or self.replica_major:
            return None
        elt = 1 if self.mode == MODE_INT8 else 4
        lat = self.n * self.NS * elt
        lat = -(-lat // 16) * 16
        RT = self.RESIDENT_THREADS
        slots = max(self.ROWS * self.BX, 1)
        chunk = min(-(-slots // RT) * RT, 16 * RT)
        lds = lat
        if self.n == 1:
            _, _, total = ising_layout(self.s)
            lds += 4 * total
        return (chunk, lds) if lds <= self.RESIDENT_LDS_MAX else None

    def emit_resident(self, plan) -> None:
        """sck_resident: one workgroup per replica keeps that replica's lattice in LDS for A.nsweep whole
        sweeps (barriers between classes replace kernel boundaries; replicas are independent chains, so no
        workgroup waits for another). Ising: a site whose table value decides it makes no draw; the others draw
        at once, so one barrier per class pass suffices (queueing them for dense draws cost a second barrier and
        was 20 % slower at the 64^2 x 512 bench shape). With several sites per thread and class, a thread reads
        the neighbours of up to RESIDENT_GROUP of them before it decides or draws any (17 % faster at that shape,
        2-6 % on 192^2 triangular and 32^3 sc lattices at B = 128)."""
        w = self.w
        T = self.elt()
        RT = self.RESIDENT_THREADS
        chunk, lds = plan
        NSn = self.n * self.NS
        w(f'extern "C" __global__ void __launch_bounds__({RT}) sck_resident(sck_args A) {{')
        w(f"    __shared__ {T} lat[{NSn}];")
        if self.n == 1:
            _, _, total = ising_layout(self.s)
            w(f"    __shared__ float tbl[{total}];")
        w("    const int rep = blockIdx.x;")
        w("    if (rep >= A.nrep) return;")
        w(f"    {T} *g = ({T} *)A.spins + (int64_t)rep * {NSn};")
        w(f"    for (int i = threadIdx.x; i < {NSn}; i += {RT}) lat[i] = g[i];")
        if self.n == 1:
            w(f"    for (int i = threadIdx.x; i < {total}; i += {RT}) tbl[i] = ((const float *)A.table)[(int64_t)rep * {total} + i];")
        w("    __syncthreads();")
        w("    for (int it = 0; it < A.nsweep; ++it) {")
        w("        sck_args Ai = A;")
        w("        Ai.sweep_offset = A.sweep_offset + it;")
        w("        const uint32_t key = sweep_key(Ai);")
        slots = self.ROWS * self.BX
        nit = -(-slots // RT)
        for k, (c, a) in enumerate(self.classes):
            w(f"        // class {k}: colour {c}, sublattice {a}")
            if self.n == 1 and nit >= 2:
                # Ising: all of a thread's sites (in groups of at most RESIDENT_GROUP) read their neighbours and
                # table value first (the LDS reads of every site in flight together), then decide or draw;
                # sites of one class are not neighbours
                nb = min(nit, RESIDENT_GROUP)
                w("        {")
                if nb < nit:
                    w(f"        for (int g0 = 0; g0 < {nit}; g0 += {nb}) {{")
                w(f"            int own_[{nb}], s0_[{nb}]; float p_[{nb}]; bool valid_[{nb}];")
                w("#pragma unroll")
                w(f"            for (int u = 0; u < {nb}; ++u) {{")
                w(f"                const int t = threadIdx.x + ({'g0 + ' if nb < nit else ''}u) * {RT};")
                w(f"                const int row = t / {self.BX};")
                w(f"                const int j = t - row * {self.BX};")
                w(f"                const int z = row / {self.L2}, y = row - (row / {self.L2}) * {self.L2};")
                w(f"                int x0 = {self.P};")
                for q in range(self.P - 1, -1, -1):
                    w(f"                if ({self.colour_expr(str(q), a)} == {c}) x0 = {q};")
                w(f"                const int x = x0 + {self.P} * j;")
                w(f"                bool valid = t < {slots} && row < {self.ROWS} && x0 < {self.P} && x < {self.L1};")
                w(f"                const int own = valid ? {a * self.PL} + row * {self.L1} + x : 0;")
                if self.has_mask:
                    w("                valid = valid && A.mask[own];")
                w("                s0_[u] = 1;")
                w(f"                p_[u] = valid ? isp_{a}(lat, 0, own, x, y, z, tbl, &s0_[u]) : 0.0f;")
                w("                own_[u] = own; valid_[u] = valid;")
                w("            }")
                # the draws of all the group's sites in one block: where a lane needs any of them, every site's
                # Philox chain runs there (a wave almost always has lanes that need each site's draw, so this
                # costs no instructions), and the independent chains interleave instead of running one branch
                # after the other
                w(f"            float ud_[{nb}];")
                w("            bool any = false;")
                w("#pragma unroll")
                w(f"            for (int u = 0; u < {nb}; ++u) any = any || (valid_[u] && !{self.decided_expr('p_[u]')});")
                w("            if (any) {")
                w("#pragma unroll")
                w(f"                for (int u = 0; u < {nb}; ++u) {{")
                w("                    const int own = own_[u];")
                self.emit_draw("                    ")
                w("                    ud_[u] = u0;")
                w("                }")
                w("            }")
                w("#pragma unroll")
                w(f"            for (int u = 0; u < {nb}; ++u) {{")
                w("                const int own = own_[u], s0 = s0_[u];")
                w("                const float p = p_[u];")
                w(f"                if (valid_[u] && {self.decided_expr('p')}) {{")
                w(f"                    if ({self.decided_value('p')}) lat[own] = ({T})(-s0);")
                w("                } else if (valid_[u]) {")
                w("                    const float u0 = ud_[u];")
                w(f"                    if ({self.accept_expr('p')}) lat[own] = ({T})(-s0);")
                w("                }")
                w("            }")
                if nb < nit:
                    w("        }")
                w("        }")
                w("        __syncthreads();")
                continue
            w(f"        for (int t = threadIdx.x; t < {-(-slots // RT) * RT}; t += {RT}) {{")
            w("            {")
            w(f"                const int row = t / {self.BX};")
            w(f"                const int j = t - row * {self.BX};")
            w(f"                const int z = row / {self.L2}, y = row - (row / {self.L2}) * {self.L2};")
            w(f"                int x0 = {self.P};")
            for q in range(self.P - 1, -1, -1):
                w(f"                if ({self.colour_expr(str(q), a)} == {c}) x0 = {q};")
            w(f"                const int x = x0 + {self.P} * j;")
            w(f"                bool valid = t < {slots} && row < {self.ROWS} && x0 < {self.P} && x < {self.L1};")
            w(f"                const int own = valid ? {a * self.PL} + row * {self.L1} + x : 0;")
            if self.has_mask:
                w("                valid = valid && A.mask[own];")
            if self.n == 1:
                # the table decides the site, or the site draws at once (the lanes that need a draw diverge)
                w("                int s0 = 1;")
                w(f"                const float p = valid ? isp_{a}(lat, 0, own, x, y, z, tbl, &s0) : 0.0f;")
                w(f"                if (valid && {self.decided_expr('p')}) {{")
                w(f"                    if ({self.decided_value('p')}) lat[own] = ({T})(-s0);")
                w("                } else if (valid) {")
                self.emit_draw("                    ")
                w(f"                    if ({self.accept_expr('p')}) lat[own] = ({T})(-s0);")
                w("                }")
            else:
                w(f"                if (valid) upd_{a}(Ai, lat, lat, 0, rep, own, x, y, z, key);")
            w("            }")
            w("        }")
            w("        __syncthreads();")
        w("    }")
        w(f"    for (int i = threadIdx.x; i < {NSn}; i += {RT}) g[i] = lat[i];")
        w("}")
        w()

    # ------------------------------------------------------------------------------ packed kernels
    def emit_packed(self) -> None:
        """Packed Ising (SCK/generator.py:157-161, 246-253, 306-308): word w of a site holds replicas
        32w .. 32w+31. Each group of 32 lanes computes the 32 replicas of one (word, site), each replica with
        its own Philox counter and beta, and the wave's ballot packs the new bits."""
        w = self.w
        s = self.s
        layouts, offsets, total = ising_layout(s)
        barker = s.acceptance == "barker"
        for a in range(self.b):
            groups, extras = _groups(s, a)
            w(f"DEV int dec_{a}(const sck_args &A, const int32_t *__restrict__ src, int64_t wbase, int lane, int rep, "
              f"int own, int x, int y, int z, uint32_t key, int *s_out) {{")
            w("    const int32_t *__restrict__ sw = src + wbase;")
            w(f"    const int s0 = (({self.word('sw', 'own', 'int32_t')} >> lane) & 1) ? 1 : -1;")
            w("    *s_out = s0;")
            gi = 0
            decl_pos = len(self.lines)
            for Jv, members in groups:
                w("    {")
                w("        int g = 0;")
                for i in members:
                    q, ok = self.nb_index(i)
                    mk = " && A.mask[q]" if self.has_mask else ""
                    w(f"        {{ const int q = {q}; int t = 0; if (({ok}){mk}) t = (({self.word('sw', 'q', 'int32_t')} >> lane) & 1) ? 1 : -1; g += t;"
                      + (f" ext{extras.index(i)} = t;" if i in extras else "") + " }")
                if Jv != 0.0:
                    w(f"        gsum{gi} = g + {len(members)};")
                    gi += 1
                w("    }")
            decl = [f"    int gsum{k} = 0;" for k in range(gi)] + [f"    int ext{k} = 0;" for k in range(len(extras))]
            self.lines[decl_pos:decl_pos] = decl
            terms, gk, ek = [], 0, 0
            for kind, size, payload in layouts[a]:
                if kind == "g":
                    terms.append((f"gsum{gk}", size)); gk += 1
                elif kind == "e":
                    terms.append((f"(ext{ek} + 1)", size)); ek += 1
                else:
                    terms.append(("(s0 > 0 ? 1 : 0)", size))
            expr = "0"
            for t, size in terms:
                expr = f"({expr}) * {size} + {t}"
            w(f"    const float p = ((const float *)A.table)[(int64_t)rep * {total} + {offsets[a]} + {expr}];")
            if barker:
                w("    if (p == 1.0f) return 1;")
                w("    if (!(p < __builtin_inff())) return 0;")
            else:
                w("    if (p >= 1.0f) return 1;")
                w("    if (!(p > 0.0f)) return 0;")
            self.emit_draw("    ")
            w("    return " + ("(u0 * p) < 1.0f;" if barker else "u0 < p;"))
            w("}")
            w()
        nsite_w = self.NS
        for k, (c, a) in enumerate(self.classes):
            w(f'extern "C" __global__ void __launch_bounds__({BLOCK}) sck_class_{k}(sck_args A) {{')
            w("    const int lane = threadIdx.x & 31;")
            w("    const int t = blockIdx.x * (blockDim.x >> 5) + (threadIdx.x >> 5);")
            w("    const int wd = blockIdx.y;")
            w(f"    const int row = t / {self.BX};")
            w(f"    const int j = t - row * {self.BX};")
            w(f"    const int z = row / {self.L2}, y = row - (row / {self.L2}) * {self.L2};")
            w(f"    bool live = row < {self.ROWS};")
            w(f"    int x0 = {self.P};")
            for q in range(self.P - 1, -1, -1):
                w(f"    if ({self.colour_expr(str(q), a)} == {c}) x0 = {q};")
            w(f"    const int x = x0 + {self.P} * j;")
            w(f"    live = live && x0 < {self.P} && x < {self.L1};")
            w(f"    const int own = live ? {a * self.PL} + row * {self.L1} + x : 0;")
            if self.has_mask:
                w("    live = live && A.mask[own];")
            w("    const uint32_t key = sweep_key(A);")
            w("    int32_t *S = (int32_t *)A.spins;")
            w(f"    const int64_t wbase = (int64_t)wd * {nsite_w};")
            w("    const int rep = wd * 32 + lane;")
            w("    int s0 = 1, acc = 0;")
            w(f"    if (live) acc = dec_{a}(A, S, wbase, lane, rep, own, x, y, z, key, &s0);")
            w("    const int nb = acc ? (s0 < 0) : (s0 > 0);")
            w("    const unsigned long long bal = __ballot(nb);")
            w("    if (live && lane == 0) S[wbase + own] = (int32_t)(uint32_t)(bal >> (threadIdx.x & 32));")
            w("}")
            w()
        # sequential: one thread per word, replicas of a word one after another at each site
        w(f'extern "C" __global__ void __launch_bounds__(64) sck_seq(sck_args A) {{')
        w("    const int wd = blockIdx.x * blockDim.x + threadIdx.x;")
        w("    if (wd >= A.nrep) return;")
        w("    const uint32_t key = sweep_key(A);")
        w("    int32_t *S = (int32_t *)A.spins;")
        w(f"    const int64_t wbase = (int64_t)wd * {nsite_w};")
        for c, a in self.classes:
            w(f"    for (int z = 0; z < {self.L3}; ++z) for (int y = 0; y < {self.L2}; ++y) for (int x = 0; x < {self.L1}; ++x) {{")
            w(f"        if ({self.colour_expr('x', a)} != {c}) continue;")
            w(f"        const int own = {a * self.PL} + (z * {self.L2} + y) * {self.L1} + x;")
            if self.has_mask:
                w("        if (!A.mask[own]) continue;")
            w("        uint32_t word = 0u;")
            w("        for (int lane = 0; lane < 32; ++lane) {")
            w(f"            int s0; const int acc = dec_{a}(A, S, wbase, lane, wd * 32 + lane, own, x, y, z, key, &s0);")
            w("            const int nb = acc ? (s0 < 0) : (s0 > 0);")
            w("            word |= (uint32_t)nb << lane;")
            w("        }")
            w("        S[wbase + own] = (int32_t)word;")
            w("    }")
        w("}")
        w()
        for a in range(self.b):
            w(f'extern "C" __global__ void __launch_bounds__({BLOCK}) sck_sync_{a}(sck_args A) {{')
            w("    const int lane = threadIdx.x & 31;")
            w("    const int t = blockIdx.x * (blockDim.x >> 5) + (threadIdx.x >> 5);")
            w("    const int wd = blockIdx.y;")
            w(f"    bool live = t < {self.PL};")
            w(f"    const int tt = live ? t : 0;")
            w(f"    const int x = tt % {self.L1}, row = tt / {self.L1};")
            w(f"    const int z = row / {self.L2}, y = row % {self.L2};")
            w(f"    const int own = {a * self.PL} + tt;")
            if self.has_mask:
                w("    live = live && A.mask[own];")
            w("    const uint32_t key = sweep_key(A);")
            w(f"    const int64_t wbase = (int64_t)wd * {nsite_w};")
            w("    int s0 = 1, acc = 0;")
            w(f"    if (live) acc = dec_{a}(A, (const int32_t *)A.spins, wbase, lane, wd * 32 + lane, own, x, y, z, key, &s0);")
            w("    const int nb = acc ? (s0 < 0) : (s0 > 0);")
            w("    const unsigned long long bal = __ballot(nb);")
            w("    if (live && lane == 0) ((int32_t *)A.dst)[wbase + own] = (int32_t)(uint32_t)(bal >> (threadIdx.x & 32));")
            w("}")
            w()

    # ----------------------------------------------------------------------------- bitwise kernels
    def bitwise_params(self):
        s = self.s
        Js = {float(s.J.get(bd.shell, 0.0)) for bd in s.crystal.bonds} - {0.0}
        h = float(s.h[0]) if s.h is not None else 0.0
        Z = max(len(self.bw_bonds(a)) for a in range(self.b))
        return Js.pop(), h, Z

    def bw_bonds(self, a):
        s = self.s
        return [i for i, bd in enumerate(s.crystal.bonds) if bd.alpha == a and s.J.get(bd.shell, 0.0) != 0.0]

    BW_ROUNDS = 4        # class slots per thread of an independent-mode block (fewer when queue items are large)

    def bw_words(self) -> int:
        """Words one thread of an independent-mode class kernel updates at each of its sites."""
        return BW_WORDS if self.W % BW_WORDS == 0 else 1

    def bw_rounds(self, a: int) -> int:
        item = 4 + 2 * max(1, self.bw_ndyn[a])
        return max(1, min(self.BW_ROUNDS, (48 * 1024) // (BLOCK * item * 4)))

    def emit_bw_indep_class(self, k: int, c: int, a: int) -> None:
        """Independent-mode class kernel with a compacted tail. Every lane of a block first runs level groups
        0 and 1 (two Philox calls: 8 of the 32 levels), which decide all 32 replicas of most (word, site)
        items; the few items still undecided are queued in LDS with their entries, and the block then runs
        the remaining level groups on dense waves of queued items. The draws and the result are those of the
        plain loop: groups are taken in order and stop only when nothing is undecided. A thread updates the
        same site in bw_words() consecutive words, so the site's row decode and neighbour offsets are shared."""
        w = self.w
        nsite = self.NS
        nd = self.bw_ndyn[a]
        WPT = self.bw_words()
        item = 4 + 2 * max(1, nd)                     # own, acc, und, e[nd], P[nd], word
        R = self.bw_rounds(a)
        qcap = max(64, R * BLOCK * WPT // BW_QUEUE_DIV)
        iw = 3 + 2 * nd
        w(f'extern "C" __global__ void __launch_bounds__({BLOCK}) sck_class_{k}(sck_args A) {{')
        w(f"    __shared__ uint32_t q[{qcap}][{item}];")
        w("    __shared__ int qn;")
        w("    if (threadIdx.x == 0) qn = 0;")
        w("    __syncthreads();")
        w(f"    const int wd0 = blockIdx.y * {WPT};")
        w("    const uint32_t key = sweep_key(A);")
        w(f"    uint32_t *S = (uint32_t *)A.spins;")
        w("    const int lane = threadIdx.x & 63;")
        w(f"    for (int r = 0; r < {R}; ++r) {{")
        w(f"        const uint32_t t = (blockIdx.x * {R} + r) * blockDim.x + threadIdx.x;")
        if self.BX % 64 == 0:
            # a wave's 64 consecutive slots lie in one row: row, y, z and the colour offset are wave-uniform
            w(f"        const int row = (int)(__builtin_amdgcn_readfirstlane(t) / {self.BX}u);")
        else:
            w(f"        const int row = (int)(t / {self.BX}u);")
        w(f"        const int j = (int)(t - (uint32_t)row * {self.BX}u);")
        w(f"        const int z = (int)((uint32_t)row / {self.L2}u), y = row - z * {self.L2};")
        w(f"        const bool rowok = row < {self.ROWS};")
        w("        bool live = rowok;")
        w(f"        int x0 = {self.P};")
        for qq in range(self.P - 1, -1, -1):
            w(f"        if ({self.colour_expr(str(qq), a)} == {c}) x0 = {qq};")
        w(f"        const int x = x0 + {self.P} * j;")
        w(f"        live = live && x0 < {self.P} && x < {self.L1};")
        w(f"        const int own = live ? {a * self.PL} + row * {self.L1} + x : 0;")
        if self.has_mask:
            w("        live = live && A.mask[own];")
        # the words' loads and neighbour counts first, before any divergent code: every word's loads keep the
        # scalar-base form and are in flight together
        w(f"        BW{a} stw[{WPT}];")
        w(f"        uint32_t s0w[{WPT}];")
        w("#pragma unroll")
        w(f"        for (int wi = 0; wi < {WPT}; ++wi) {{")
        # y and z stay in range whenever the row is (a dead lane's x is zeroed): with a wave-uniform row the
        # coordinates and their neighbour offsets stay scalar
        w(f"            stw[wi] = setup_{a}(A, S, wd0 + wi, own, live ? x : 0, rowok ? y : 0, rowok ? z : 0, &s0w[wi]);")
        w("            if (!live) { stw[wi].acc = 0u; stw[wi].und = 0u; }")
        w("        }")
        w("#pragma unroll")
        w(f"        for (int wi = 0; wi < {WPT}; ++wi) {{")
        w("            const int wd = wd0 + wi;")
        w(f"            uint32_t *Sw = S + (int64_t)wd * {nsite};")
        w("            const uint32_t s0 = s0w[wi];")
        w(f"            BW{a} &st = stw[wi];")
        w(f"            const uint32_t ctr = (uint32_t)wd * {nsite}u + (uint32_t)own;")
        w(f"            levels_{a}(st, ctr, key, 0, 2);")
        w("            const bool tail = st.und != 0u;")
        w(f"            if (live && !tail) {self.word('Sw', 'own', const=False)} = s0 ^ st.acc;")
        w("            const unsigned long long m = __ballot(tail);")
        w("            if (m) {")
        w("                int qb = 0;")
        w("                if (lane == 0) qb = atomicAdd(&qn, __popcll(m));")
        w("                qb = __shfl(qb, 0, 64);")
        w("                if (tail) {")
        w("                    const int pos = qb + __popcll(m & ((1ull << lane) - 1ull));")
        w(f"                    if (pos < {qcap}) {{")
        w(f"                        q[pos][0] = (uint32_t)own; q[pos][1] = st.acc; q[pos][2] = st.und; "
          f"q[pos][{iw}] = (uint32_t)wd;")
        for i in range(nd):
            w(f"                        q[pos][{3 + i}] = st.e[{i}]; q[pos][{3 + nd + i}] = st.P[{i}];")
        w("                    } else {")
        w("                        // queue full: finish the item here, with the same level groups in the same order")
        w(f"                        levels_{a}(st, ctr, key, 2, {LEVELS // 4});")
        w(f"                        {self.word('Sw', 'own', const=False)} = s0 ^ st.acc;")
        w("                    }")
        w("                }")
        w("            }")
        w("        }")
        w("    }")
        w("    __syncthreads();")
        w(f"    const int n = qn < {qcap} ? qn : {qcap};")
        w(f"    for (int i = threadIdx.x; i < ((n + 63) & ~63); i += {BLOCK}) {{")
        w("        const bool has = i < n;")
        w(f"        BW{a} st;")
        w("        const int own = has ? (int)q[i][0] : 0;")
        w(f"        const int wd = has ? (int)q[i][{iw}] : wd0;")
        w("        st.acc = has ? q[i][1] : 0u; st.und = has ? q[i][2] : 0u;")
        for i in range(nd):
            w(f"        st.e[{i}] = has ? q[i][{3 + i}] : 0u; st.P[{i}] = has ? q[i][{3 + nd + i}] : 0u;")
        w(f"        levels_{a}(st, (uint32_t)wd * {nsite}u + (uint32_t)own, key, 2, {LEVELS // 4});")
        w(f"        uint32_t *Sw = S + (int64_t)wd * {nsite};")
        w(f"        if (has) {self.word('Sw', 'own', const=False)} ^= st.acc;")
        w("    }")
        w("}")
        w()

    def emit_count_planes(self, xs: list, NP: int) -> None:
        """c0 .. c{NP-1}: the bit-planes of the per-lane count of set bits across the planes `xs` (count <
        2^NP), by a carry-save tree of full adders (3-input xor and majority, one bitop3 each) and half adders,
        column by column from the lowest weight."""
        w = self.w
        cols = [list(xs)] + [[] for _ in range(NP)]
        tmp = 0
        for wt in range(NP):
            col = cols[wt]
            while len(col) >= 3:
                a, b, c = col.pop(0), col.pop(0), col.pop(0)
                sname, cname = f"f{tmp}s", f"f{tmp}c"
                tmp += 1
                # truth tables over (a, b, c) = (0xF0, 0xCC, 0xAA): a ^ b ^ c = 0x96, majority = 0xE8
                w(f"    const uint32_t {sname} = __builtin_amdgcn_bitop3_b32({a}, {b}, {c}, 0x96);")
                if wt + 1 < NP:
                    w(f"    const uint32_t {cname} = __builtin_amdgcn_bitop3_b32({a}, {b}, {c}, 0xE8);")
                    cols[wt + 1].append(cname)
                col.append(sname)
            if len(col) == 2:
                a, b = col.pop(0), col.pop(0)
                sname, cname = f"f{tmp}s", f"f{tmp}c"
                tmp += 1
                w(f"    const uint32_t {sname} = {a} ^ {b};")
                if wt + 1 < NP:
                    w(f"    const uint32_t {cname} = {a} & {b};")
                    cols[wt + 1].append(cname)
                col.append(sname)
            # the count is below 2^NP, so no carry leaves the top plane
            w(f"    const uint32_t c{wt} = {col[0] if col else '0u'};")

    def bw_entries(self, a: int):
        """Flip-mask entries of the independent mode, one per k (and per spin sign with a field), as
        (kind, lane-mask expression, table offset): 'static' entries (Metropolis on a periodic unmasked box
        with dE <= 0, p = 1) accept with no draw (SCK/bitwise.py:373-384), 'dyn' entries compare U with P."""
        s = self.s
        J, h, Z = self.bitwise_params()
        Z1 = Z + 1
        field = h != 0.0
        static = s.acceptance != "barker" and s.mask is None and all(s.periodic)
        out = []
        for k in range(Z1):
            sels = ([(f"(eq{k}) & ~s0", k, -1), (f"(eq{k}) & s0", Z1 * Z1 + k, 1)] if field else [(f"eq{k}", k, -1)])
            for sel, off, sigma in sels:
                if static and k <= Z and 2.0 * J * (Z - 2 * k) + 2.0 * h * sigma <= 0.0:
                    out.append(("static", sel, off))
                else:
                    out.append(("dyn", sel, off))
        return out

    def emit_bitwise(self) -> None:
        """Bitwise multi-spin Ising (SCK/bitwise.py:250-422). A lane owns one site of one word (32 replicas):
        anti-aligned neighbour counts go into NP bit planes through a ripple adder, and the flip mask is
        built from the table p[word, s, z_eff, k]. Shared mode: one uniform per (word, site), output w % 4 of
        the Philox call with counter (w / 4) nsite + i. Independent mode: each replica's 32-bit U is compared
        with P = floor(p 2^32) bit-sliced, most significant level first; four levels per Philox call with the
        key of the level group, counter w nsite + i; a wave stops drawing once all its lanes are decided."""
        w = self.w
        s = self.s
        J, h, Z = self.bitwise_params()
        NP = max(1, Z.bit_length())
        Z1 = Z + 1
        NT = 2 * Z1 * Z1
        field = h != 0.0
        indep = self.mode == MODE_BITWISE_INDEP
        static = (indep and s.acceptance != "barker" and s.mask is None and all(s.periodic))
        nsite = self.NS
        self.bw_ne, self.bw_ndyn = {}, {}
        for a in range(self.b):
            if indep:
                nd = sum(1 for kind, _, _ in self.bw_entries(a) if kind != "static")
                self.bw_ndyn[a] = nd
                w(f"struct BW{a} {{ uint32_t acc, und; uint32_t e[{max(1, nd)}], P[{max(1, nd)}]; }};")
                w(f"DEV BW{a} setup_{a}(const sck_args &A, const uint32_t *__restrict__ S, int wd, int own, int x, int y, "
                  f"int z, uint32_t *s_out) {{")
            else:
                w(f"DEV uint32_t flip_{a}(const sck_args &A, const uint32_t *__restrict__ S, int wd, int own, int x, int y, "
                  f"int z, uint32_t key) {{")
            w(f"    const uint32_t *__restrict__ Sw = S + (int64_t)wd * {nsite};")
            w(f"    const uint32_t s0 = {self.word('Sw', 'own')};")
            if indep:
                w("    *s_out = s0;")
            w("    int ze = 0;")
            # anti-aligned present neighbours, one bit-plane each; their count goes into NP bit-planes c_p
            xs = []
            for j, i in enumerate(self.bw_bonds(a)):
                q, ok = self.nb_index(i)
                mk = " && A.mask[q]" if self.has_mask else ""
                if ok == "true" and not mk:
                    w(f"    ze += 1; const uint32_t x{j} = s0 ^ {self.word('Sw', f'({q})')};")
                else:
                    w(f"    uint32_t x{j} = 0u; {{ const int q = {q}; if (({ok}){mk}) {{ ze += 1; "
                      f"x{j} = s0 ^ {self.word('Sw', 'q')}; }} }}")
                xs.append(f"x{j}")
            self.emit_count_planes(xs, NP)
            eqs = []
            for k in range(Z1):
                terms = [f"c{p}" if (k >> p) & 1 else f"~c{p}" for p in range(NP)]
                eqs.append(" & ".join(terms))
            if not indep:
                w(f"    const float *tb = (const float *)A.table + (int64_t)wd * {NT} + ze * {Z1};")
                w(f"    uint32_t c0r = (uint32_t)(wd >> 2) * {nsite}u + (uint32_t)own, c1r = 0u, c2r = 0u, c3r = 0u;")
                w("    philox10_z(c0r, c1r, c2r, c3r, key);")
                w("    const int qsel = wd & 3;")
                w("    const uint32_t raw = qsel == 0 ? c0r : (qsel == 1 ? c1r : (qsel == 2 ? c2r : c3r));")
                w("    const float u = u01(raw);")
                w("    uint32_t acc = 0u;")
                for k in range(Z1):
                    w(f"    {{ const uint32_t eq = {eqs[k]};")
                    if field:
                        w(f"      if (u < tb[{Z1 * Z1 + k}]) acc |= eq & s0;")
                        w(f"      if (u < tb[{k}]) acc |= eq & ~s0; }}")
                    else:
                        w(f"      if (u < tb[{k}]) acc |= eq; }}")
                w("    return acc;")
                w("}")
                w()
                continue
            # independent mode: setup (entries and their P), then the level loop
            w(f"    const int64_t *tb = (const int64_t *)A.table + (int64_t)wd * {NT} + ze * {Z1};")
            ent = self.bw_entries(a)
            self.bw_ne[a] = len(ent)
            w(f"    BW{a} st;")
            w("    st.acc = 0u; st.und = 0u;")
            for k in range(Z1):
                w(f"    const uint32_t eq{k} = {eqs[k]};")
            for kind, sel, off in ent:
                if kind == "static":
                    w(f"    st.acc |= {sel};")
                    continue
            ei = 0
            for kind, sel, off in ent:
                if kind == "static":
                    continue
                w(f"    {{ const int64_t Pq = tb[{off}]; st.e[{ei}] = {sel}; st.P[{ei}] = (uint32_t)Pq;")
                w(f"      if (Pq >= 4294967296LL) st.acc |= st.e[{ei}];")
                w(f"      if (Pq > 0 && Pq < 4294967296LL) st.und |= st.e[{ei}]; }}")
                ei += 1
            w("    return st;")
            w("}")
            w()
            nd = self.bw_ndyn[a]
            # level groups [g0, g1): four levels per Philox call; a wave skips a group nobody needs
            w(f"DEV void levels_{a}(BW{a} &st, uint32_t ctr, uint32_t key, int g0, int g1) {{")
            if nd:
                w("    for (int g = g0; g < g1; ++g) {")
                w("        if (__ballot(st.und != 0u) == 0ull) break;")
                # (key + (g + 1) * step) mod 2147483629 with key already reduced and (g + 1) * step < 2^20: one
                # conditional subtract reduces it (scalar work per level group)
                w(f"        const uint32_t kx = key + (uint32_t)(g + 1) * {GROUP_KEY_STEP}u;")
                w("        const uint32_t kg = kx >= 2147483629u ? kx - 2147483629u : kx;")
                w("        uint32_t r0 = ctr, r1 = 0u, r2 = 0u, r3 = 0u;")
                w("        philox10_z(r0, r1, r2, r3, kg);")
                w("        const uint32_t rw[4] = {r0, r1, r2, r3};")
                # level 4g + q reads bit 31 - 4g - q of each P: shift the group's bits to the top once,
                # then spread bit 31 - q into an all-ones or all-zero mask (uniform P: scalar work)
                for i in range(nd):
                    w(f"        const uint32_t Pg{i} = st.P[{i}] << (4 * g);")
                w("#pragma unroll")
                w("        for (int q = 0; q < 4; ++q) {")
                # bitop3 truth tables over (a, b, c) = (0xF0, 0xCC, 0xAA): (a & b) | c = 0xEA,
                # a & ~b & c = 0x20, a & ~(b ^ c) = 0x90
                w("            uint32_t pb = st.e[0] & (uint32_t)((int32_t)(Pg0 << q) >> 31);")
                for i in range(1, nd):
                    w(f"            pb = __builtin_amdgcn_bitop3_b32(st.e[{i}], (uint32_t)((int32_t)(Pg{i} << q) >> 31), "
                      f"pb, 0xEA);")
                w("            st.acc |= __builtin_amdgcn_bitop3_b32(st.und, rw[q], pb, 0x20);")
                w("            st.und = __builtin_amdgcn_bitop3_b32(st.und, rw[q], pb, 0x90);")
                w("        }")
                w("    }")
            w("}")
            w()
            w(f"DEV uint32_t flip_{a}(const sck_args &A, const uint32_t *__restrict__ S, int wd, int own, int x, int y, "
              f"int z, uint32_t key) {{")
            w(f"    uint32_t s0; BW{a} st = setup_{a}(A, S, wd, own, x, y, z, &s0);")
            w(f"    levels_{a}(st, (uint32_t)wd * {nsite}u + (uint32_t)own, key, 0, {LEVELS // 4});")
            w("    return st.acc;")
            w("}")
            w()
        # class kernels: lane = site of the class row tile, blockIdx.y = word
        for k, (c, a) in enumerate(self.classes):
            if indep:
                self.emit_bw_indep_class(k, c, a)
                continue
            w(f'extern "C" __global__ void __launch_bounds__({BLOCK}) sck_class_{k}(sck_args A) {{')
            w("    const int t = blockIdx.x * blockDim.x + threadIdx.x;")
            w("    const int wd = blockIdx.y;")
            w(f"    const int row = t / {self.BX};")
            w(f"    const int j = t - row * {self.BX};")
            w(f"    const int z = row / {self.L2}, y = row - (row / {self.L2}) * {self.L2};")
            w(f"    bool live = row < {self.ROWS};")
            w(f"    int x0 = {self.P};")
            for q in range(self.P - 1, -1, -1):
                w(f"    if ({self.colour_expr(str(q), a)} == {c}) x0 = {q};")
            w(f"    const int x = x0 + {self.P} * j;")
            w(f"    live = live && x0 < {self.P} && x < {self.L1};")
            w(f"    const int own = live ? {a * self.PL} + row * {self.L1} + x : 0;")
            if self.has_mask:
                w("    live = live && A.mask[own];")
            if indep:
                raise AssertionError("independent class kernels are emitted by emit_bw_indep_class")
            else:
                w("    if (!live) return;")
                w("    uint32_t *S = (uint32_t *)A.spins;")
                w("    const uint32_t key = sweep_key(A);")
                w(f"    uint32_t *Sw = S + (int64_t)wd * {nsite};")
                w(f"    {self.word('Sw', 'own', const=False)} ^= flip_{a}(A, S, wd, own, x, y, z, key);")
            w("}")
            w()
        w(f'extern "C" __global__ void __launch_bounds__(64) sck_seq(sck_args A) {{')
        w("    const int wd = blockIdx.x * blockDim.x + threadIdx.x;")
        w("    if (wd >= A.nrep) return;")
        w("    const uint32_t key = sweep_key(A);")
        w("    uint32_t *S = (uint32_t *)A.spins;")
        for c, a in self.classes:
            w(f"    for (int z = 0; z < {self.L3}; ++z) for (int y = 0; y < {self.L2}; ++y) for (int x = 0; x < {self.L1}; ++x) {{")
            w(f"        if ({self.colour_expr('x', a)} != {c}) continue;")
            w(f"        const int own = {a * self.PL} + (z * {self.L2} + y) * {self.L1} + x;")
            if self.has_mask:
                w("        if (!A.mask[own]) continue;")
            w(f"        S[(int64_t)wd * {nsite} + own] ^= flip_{a}(A, S, wd, own, x, y, z, key);")
            w("    }")
        w("}")
        w()
        for a in range(self.b):
            w(f'extern "C" __global__ void __launch_bounds__({BLOCK}) sck_sync_{a}(sck_args A) {{')
            w("    const int t = blockIdx.x * blockDim.x + threadIdx.x;")
            w("    const int wd = blockIdx.y;")
            w(f"    bool live = t < {self.PL};")
            w("    const int tt = live ? t : 0;")
            w(f"    const int x = tt % {self.L1}, row = tt / {self.L1};")
            w(f"    const int z = row / {self.L2}, y = row % {self.L2};")
            w(f"    const int own = {a * self.PL} + tt;")
            if self.has_mask:
                w("    live = live && A.mask[own];")
            w("    const uint32_t key = sweep_key(A);")
            w(f"    const uint32_t f = flip_{a}(A, (const uint32_t *)A.spins, wd, own, x, y, z, key);")
            w(f"    if (live) ((uint32_t *)A.dst)[(int64_t)wd * {nsite} + own] = ((const uint32_t *)A.spins)[(int64_t)wd * {nsite} + own] ^ f;")
            w("}")
            w()

    # --------------------------------------------------------------------------------- observables
    def emit_observe(self, chunk: int) -> None:
        """Float64 energy and per-sublattice magnetization (SCK/sampler.py:296-345): each block sums the
        sites [blockIdx.x * chunk, ...) of its replicas into one partial per replica; sck_obs_final adds the
        partials of a replica in a fixed order.

        Replica-major: a workgroup is 64 replicas (the lanes) and 4 waves, each wave a quarter of the chunk's
        sites; the waves' sums are added in LDS in wave order. A launch covers the replicas [A.rep_base,
        A.rep_base + obs_span) and its partial sums sit at rows rep - A.rep_base of A.obs_scratch, so the
        observation of a replica-blocked module runs block by block (each block's observation right after its
        sweeps, while its lattice is cache-resident) and reuses one block's rows of scratch."""
        w, s, n, b = self.w, self.s, self.n, self.b
        packed = self.mode in (MODE_PACKED, MODE_BITWISE, MODE_BITWISE_INDEP)
        NV = 1 + b * n
        w("DEV double spin_at(const sck_args &A, int rep, int comp, int q) {")
        if packed:
            w(f"    const uint32_t wv = ((const uint32_t *)A.spins)[(int64_t)(rep >> 5) * {self.NS} + q];")
            w("    const double sv = ((wv >> (rep & 31)) & 1u) ? 1.0 : -1.0;")
            w("    return A.mask[q] ? sv : 0.0;")
        else:
            T = self.elt()
            if self.replica_major:
                w(f"    return (double)((const {T} *)A.spins)[((int64_t)q * {self.n} + comp) * {self.B} + rep];")
            else:
                w(f"    return (double)((const {T} *)A.spins)[(int64_t)rep * {self.n * self.NS} + (int64_t)comp * {self.NS} + q];")
        w("}")
        w()
        w(f'extern "C" __global__ void __launch_bounds__({BLOCK}) sck_obs(sck_args A) {{')
        if self.replica_major:
            # lanes over replicas (uniform site addresses, one 256-byte access per spin component); the 4 waves
            # split the chunk's sites
            sub = -(-chunk // (BLOCK // 64))
            w("    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;")
            w("    const int rel = blockIdx.y * 64 + lane;")
            w("    const int rep = A.rep_base + rel;")
            w(f"    const bool live = rel < {self.obs_span} && rep < {self.B};")
            w(f"    double acc[{NV}];")
            w(f"    for (int k = 0; k < {NV}; ++k) acc[k] = 0.0;")
            w(f"    const int lo = blockIdx.x * {chunk} + wave * {sub};")
            w(f"    const int hi = min(min(lo + {sub}, (int)(blockIdx.x + 1) * {chunk}), {self.NS});")
            w("    for (int site = lo; live && site < hi; ++site) {")
        else:
            w("    const int rep = blockIdx.y;")
            w(f"    double acc[{NV}];")
            w(f"    for (int k = 0; k < {NV}; ++k) acc[k] = 0.0;")
            w(f"    const int lo = blockIdx.x * {chunk};")
            w(f"    const int hi = min(lo + {chunk}, {self.NS});")
            w("    for (int site = lo + threadIdx.x; site < hi; site += blockDim.x) {")
        self.emit_site_energy(lambda k, q: f"spin_at(A, rep, {k}, {q})")
        w("    }")
        if self.replica_major:
            W = BLOCK // 64
            w(f"    __shared__ double red[{W}][{NV}][64];")
            w(f"    for (int k = 0; k < {NV}; ++k) red[wave][k][lane] = acc[k];")
            w("    __syncthreads();")
            w("    if (wave == 0 && live) {")
            w(f"        for (int k = 0; k < {NV}; ++k) {{")
            w("            double v = red[0][k][lane];")
            w(f"            for (int i = 1; i < {W}; ++i) v += red[i][k][lane];")
            w(f"            A.obs_scratch[((int64_t)rel * gridDim.x + blockIdx.x) * {NV} + k] = v;")
            w("        }")
            w("    }")
            w("}")
        else:
            self.emit_obs_block_reduce(NV)
        self.emit_obs_final(NV)

    def energy_bonds(self) -> dict:
        """Bonds the observables sum, with their weight. The reference halves the sum over every directed bond
        (SCK/sampler.py:326-341). Where every bond (a -> a2, delta) has its reverse (a2 -> a, -delta) with the
        transposed matrix, summing one bond of each pair at weight 1 gives the same energy (the reverse bond's
        term at a site is the forward bond's term at its neighbour, and a box edge or a mask drops both), with
        half the neighbour reads; otherwise every bond is summed at weight 1/2."""
        s = self.s
        bonds = s.crystal.bonds
        index = {(bd.alpha, bd.alpha2, tuple(bd.delta)): i for i, bd in enumerate(bonds)}
        if len(index) != len(bonds):
            return {i: 0.5 for i in range(len(bonds))}
        keep = {}
        paired = set()
        for i, bd in enumerate(bonds):
            j = index.get((bd.alpha2, bd.alpha, tuple(-d for d in bd.delta)))
            if j is None or j == i or not np.array_equal(s.bond_matrix(j), s.bond_matrix(i).T):
                return {i: 0.5 for i in range(len(bonds))}
            if i not in paired:
                keep[i] = 1.0
                paired.add(j)
        return keep

    def emit_site_energy(self, spin) -> None:
        """Loop body of the observables: site `site`'s coordinates, its spins s<k>, and its share of the energy
        (bond terms halved) and magnetization added to acc[]; spin(k, q) is the expression of component k of
        site q of the replica being summed."""
        w, s, n, b = self.w, self.s, self.n, self.b
        w(f"        const int a = site / {self.PL}, rem = site - a * {self.PL};")
        w(f"        const int x = rem % {self.L1}, row = rem / {self.L1};")
        w(f"        const int z = row / {self.L2}, y = row % {self.L2};")
        for k in range(n):
            w(f"        const double s{k} = {spin(k, 'site')};")
        w("        double e = 0.0;")
        weights = self.energy_bonds()
        for a in range(b):
            w(f"        if (a == {a}) {{")
            for k in range(n):
                w(f"            acc[{1 + a * n + k}] += s{k};")
            for i, bd in enumerate(s.crystal.bonds):
                if bd.alpha != a or i not in weights:
                    continue
                M = s.bond_matrix(i)
                if not M.any():
                    continue
                q, ok = self.nb_index(i)
                half = "0.5 * " if weights[i] == 0.5 else ""
                w(f"            if ({ok}) {{ const int q = {q};")
                for l in range(n):
                    w(f"                const double n{l} = {spin(l, 'q')};")
                for k in range(n):
                    terms = [f"{_f64(M[k, l])} * n{l}" for l in range(n) if M[k, l] != 0.0]
                    if terms:
                        w(f"                e -= {half}s{k} * ({' + '.join(terms)});")
                w("            }")
            w("        }")
        if s.h is not None:
            for k in range(n):
                if s.h[k] != 0.0:
                    w(f"        e -= {_f64(s.h[k])} * s{k};")
        if s.cubic_k1 is not None:
            w(f"        e += (double)A.k1[site] * ((s0 * s1) * (s0 * s1) + (s0 * s2) * (s0 * s2) + (s1 * s2) * (s1 * s2))"
              f" + {_f64(s.cubic_k2)} * (s0 * s1 * s2) * (s0 * s1 * s2);")
        if s.uniaxial is not None:
            Du, ax = s.uniaxial
            se = " + ".join(f"s{k} * {_f64(ax[k])}" for k in range(n))
            w(f"        {{ const double se = {se}; e -= {_f64(Du)} * se * se; }}")
        if s.site_ku is not None:
            sn = " + ".join(f"s{k} * (double)A.site_axis[(int64_t)site * {n} + {k}]" for k in range(n))
            w(f"        {{ const double sn = {sn}; e -= (double)A.site_ku[site] * sn * sn; }}")
        w("        acc[0] += e;")

    def emit_obs_block_reduce(self, NV: int) -> None:
        w = self.w
        # block reduction: wave butterfly, then LDS across the waves
        w(f"    __shared__ double red[{BLOCK // 64}][{NV}];")
        w(f"    for (int k = 0; k < {NV}; ++k) {{")
        w("        double v = acc[k];")
        w("        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);")
        w("        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = v;")
        w("    }")
        w("    __syncthreads();")
        w(f"    if (threadIdx.x < {NV}) {{")
        w("        double v = 0.0;")
        w(f"        for (int i = 0; i < {BLOCK // 64}; ++i) v += red[i][threadIdx.x];")
        w(f"        A.obs_scratch[((int64_t)rep * gridDim.x + blockIdx.x) * {NV} + threadIdx.x] = v;")
        w("    }")
        w("}")

    def emit_obs_final(self, NV: int) -> None:
        w = self.w
        w()
        if self.replica_major:
            # a wave per replica: lane l adds partials l, l + 64, ... in order, then a butterfly over the lanes
            # (the same association every run, so the result is deterministic)
            w(f'extern "C" __global__ void __launch_bounds__({BLOCK}) sck_obs_final(sck_args A) {{')
            w("    const int lane = threadIdx.x & 63;")
            w(f"    const int rel = blockIdx.x * {BLOCK // 64} + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);")
            w("    const int rep = A.rep_base + rel;")
            w(f"    if (rel >= {self.obs_span} || rep >= {self.B}) return;")
            w(f"    const int nch = {self.obs_chunks};")
            w(f"    for (int k = 0; k < {NV}; ++k) {{")
            w("        double v = 0.0;")
            w(f"        for (int c = lane; c < nch; c += 64) v += A.obs_scratch[((int64_t)rel * nch + c) * {NV} + k];")
            w("        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);")
            w("        if (lane == 0) {")
            w("            if (k == 0) A.obs[rep] = v;")
            w(f"            else A.obs[{self.B} + (int64_t)rep * {NV - 1} + (k - 1)] = v;")
            w("        }")
            w("    }")
            w("}")
        else:
            w(f'extern "C" __global__ void __launch_bounds__(64) sck_obs_final(sck_args A) {{')
            w("    const int rep = blockIdx.x * blockDim.x + threadIdx.x;")
            w(f"    if (rep >= {self.B}) return;")
            w(f"    const int nch = {self.obs_chunks};")
            w(f"    for (int k = 0; k < {NV}; ++k) {{")
            w("        double v = 0.0;")
            w(f"        for (int c = 0; c < nch; ++c) v += A.obs_scratch[((int64_t)rep * nch + c) * {NV} + k];")
            w("        if (k == 0) A.obs[rep] = v;")
            w(f"        else A.obs[{self.B} + (int64_t)rep * {NV - 1} + (k - 1)] = v;")
            w("    }")
            w("}")
        w()
        w('extern "C" __global__ void sck_bump(sck_args A) {')
        w("    if (threadIdx.x == 0 && blockIdx.x == 0) ((int64_t *)A.state)[1] += A.sweep_offset;")
        w("}")

    # ------------------------------------------------------------------------------------- driver
    def generate(self) -> Module:
        s = self.s
        self.w(HEADER)
        self.w("#pragma clang fp contract(off)")
        self.w(f"// system: {s.name}; n={s.n} b={s.b} box={s.box} periodic={s.periodic}; "
               f"colouring {s.colouring.describe()}; mode {self.mode}; proposal {self.proposal}")
        self.w()
        rplan = self.resident_plan()
        tplan = self.tiled_plan() if self.want_tiled else None
        if self.want_tiled and tplan is None:
            raise ValueError("the tiled sweep needs float32 vector spins on a 2-D lattice in the reference layout, "
                             "larger than a tile and its halo")
        if self.mode in (MODE_INT8, MODE_F32, MODE_F64):
            self.emit_plain()
            if rplan is not None:
                self.emit_resident(rplan)
            if tplan is not None:
                self.emit_tiled(tplan)
        elif self.mode == MODE_PACKED:
            self.emit_packed()
        else:
            self.emit_bitwise()
        # observables: enough blocks per replica for a full chip when replicas are few
        self.obs_span = self.B
        if self.replica_major:
            # one launch per replica block: about 2048 workgroups of 64 replicas x 4 waves each
            self.obs_span = self.rm_block
            groups = -(-self.obs_span // 64)
            per = max(1, min(-(-self.NS // (BLOCK // 64)), -(-2048 // groups)))
            chunk = -(-self.NS // per)
        else:
            per = max(1, min(64, (4096 + self.B - 1) // max(1, self.B)))
            chunk = max(BLOCK, -(-self.NS // per))
            chunk = -(-chunk // BLOCK) * BLOCK
        self.obs_chunks = -(-self.NS // chunk)
        self.emit_observe(chunk)
        src = "\n".join(self.lines) + "\n"
        kernels = [f"sck_class_{k}" for k in range(len(self.classes))] + ["sck_seq"] + \
                  [f"sck_sync_{a}" for a in range(self.b)] + ["sck_obs", "sck_obs_final", "sck_bump"] + \
                  (["sck_resident"] if rplan is not None else []) + \
                  (["sck_tiled", "sck_track_fold"] if tplan is not None else []) + \
                  (["sck_sample"] if (self.n > 1 and self.mode in (MODE_F32, MODE_F64)) else [])
        ki = {k: i for i, k in enumerate(kernels)}
        R = self.W if self.mode in (MODE_PACKED, MODE_BITWISE, MODE_BITWISE_INDEP) else self.B
        gy, gz = (R, 1) if R <= 65535 else (65535, -(-R // 65535))
        sweep = []
        nblk = self.B // self.rm_block if self.replica_major else 1
        for k in range(len(self.classes)):
            if self.mode == MODE_PACKED:
                sweep.append(Launch(ki[f"sck_class_{k}"], (-(-self.ROWS * self.BX // (BLOCK // 32)), self.W, 1)))
            elif self.mode == MODE_BITWISE_INDEP:
                tile = BLOCK * self.bw_rounds(self.classes[k][1])
                sweep.append(Launch(ki[f"sck_class_{k}"], (-(-self.ROWS * self.BX // tile),
                                                            self.W // self.bw_words(), 1)))
            elif self.replica_major:
                _, _, chunk = self.rm_plan()
                gy = min(chunk, 65535)
                sweep.append(Launch(ki[f"sck_class_{k}"], (XCDS, gy, -(-chunk // gy))))
            else:
                sweep.append(Launch(ki[f"sck_class_{k}"], (-(-self.ROWS * self.BX // BLOCK), gy, gz)))
        ref = [Launch(ki["sck_seq"], (-(-R // 64), 1, 1), (64, 1, 1))]
        sync = []
        for a in range(self.b):
            if self.mode == MODE_PACKED:
                sync.append(Launch(ki[f"sck_sync_{a}"], (-(-self.PL // (BLOCK // 32)), self.W, 1)))
            elif self.replica_major:
                gy2 = min(self.PL, 65535)
                sync.append(Launch(ki[f"sck_sync_{a}"], (-(-self.B // BLOCK), gy2, -(-self.PL // gy2))))
            else:
                sync.append(Launch(ki[f"sck_sync_{a}"], (-(-self.PL // BLOCK), gy, gz)))
        if self.replica_major:
            # one (partial sums, fold) pair per replica block, each block's sums in the same scratch rows
            span = self.obs_span
            obs = []
            for rb in range(nblk):
                obs.append(Launch(ki["sck_obs"], (self.obs_chunks, -(-span // 64), 1), rep_base=rb * span))
                obs.append(Launch(ki["sck_obs_final"], (-(-span // (BLOCK // 64)), 1, 1), rep_base=rb * span))
        else:
            obs = [Launch(ki["sck_obs"], (self.obs_chunks, self.B, 1)),
                   Launch(ki["sck_obs_final"], (-(-self.B // 64), 1, 1), (64, 1, 1))]
        bump = Launch(ki["sck_bump"], (1, 1, 1), (64, 1, 1))
        layouts, offsets, total = ising_layout(s) if s.n == 1 else ([], [], 0)
        if nblk > 1:
            # block-major: the class launches of replica block 0, then of block 1, ...
            sweep = [replace(l, rep_base=rb * self.rm_block) for rb in range(nblk) for l in sweep]
        mod = Module(src, kernels, sweep, ref, sync, obs, bump, self.obs_chunks, self.classes,
                     table_entries=total, ising_layout=layouts, rep_blocks=nblk,
                     obs_blocks=nblk if self.replica_major else 1, obs_rows=self.obs_span)
        if rplan is not None:
            mod.resident = Launch(ki["sck_resident"], (self.B, 1, 1), (self.RESIDENT_THREADS, 1, 1))
        if tplan is not None:
            # persistent: grid.y = 0 asks the launcher for the workgroups resident per XCD (from the occupancy)
            mod.tiled = Launch(ki["sck_tiled"], (XCDS, 0, 1), (self.TILE_THREADS, 1, 1))
            mod.tile_dims = tuple(tplan[:2]) + (self.TILE_THREADS,)
            nfold = self.B * (1 + self.b * self.n)
            mod.track_fold = Launch(ki["sck_track_fold"], (-(-nfold // 256), 1, 1), (256, 1, 1))
            mod.track_parts = self.tiled_nt * nfold     # doubles of one sweep's partial sums
        mod.key = hashlib.sha256(src.encode()).hexdigest()[:16]
        return mod


# ------------------------------------------------------------------------------------ batched systems
BATCH_HDR = 4                 # int64 header of the batch table: systems, site-list offset (uint32 words), classes, 0
CLASS_BOUND = 256             # launch bound of the batch class kernels (schedules use 64 .. 256 threads)
CHAIN_BLOCKS = (256, 1024)    # workgroup sizes of the chain kernels
COORD_BITS = 10               # packed site coordinates x | y << 10 | z << 20


def batch_stride(nclass: int) -> int:
    """int64 words of one system's descriptor: L1, L2, L3, replicas, first global replica, Philox seed, spin
    offset (elements), site offset, then (first, count) of each class's site list."""
    return 8 + 2 * nclass


@dataclass
class BatchModule:
    src: str
    kernels: list[str]
    classes: list[tuple[int, int]]
    table_entries: int = 0
    key: str = ""


class _BatchGen(_Gen):
    """Kernels for a batch of systems that share one Hamiltonian structure (model, crystal, couplings, extra
    matrices, field, anisotropy form, colouring classes) but differ in box, mask and per-site constants. The
    box is a runtime value read from the system's descriptor; each system keeps its own Philox seed, its own
    counter layout rep * nsite + own and its own spin array in the reference layout, so a system in the batch
    runs exactly the chain it runs alone."""

    def __init__(self, systems, mode, proposal, learned_meta, wide):
        super().__init__(systems[0], mode, 1, proposal, learned_meta, replica_major=False)
        esz = 1 if mode == MODE_INT8 else 4
        self.u32 = max(s.n * s.nsite for s in systems) * esz < 2 ** 31
        self.rt = True
        self.L1, self.L2, self.L3, self.PL, self.NS = "D.L1", "D.L2", "D.L3", "D.PL", "D.NS"
        self.wide = bool(wide)
        self.T = self.elt()
        self.stride = batch_stride(len(self.classes))

    def emit_batch_helpers(self) -> None:
        w, n, T = self.w, self.n, self.T
        _, _, total = ising_layout(self.s) if n == 1 else ([], [], 0)
        w("typedef struct bdims { int L1, L2, L3, PL, NS; } bdims;")
        w()
        w("DEV const int64_t *bdesc(const sck_args &G, int s) {")
        w(f"    return (const int64_t *)G.aux + {BATCH_HDR} + (int64_t)s * {self.stride};")
        w("}")
        w()
        w("DEV const uint32_t *blist(const sck_args &G) {")
        w("    return (const uint32_t *)G.aux + ((const int64_t *)G.aux)[1];")
        w("}")
        w()
        w("// the system's box and its arguments: pointers moved to its spins, site constants and first replica")
        w("DEV void bsys(const sck_args &G, const int64_t *d, bdims &D, sck_args &A) {")
        w("    D.L1 = (int)d[0]; D.L2 = (int)d[1]; D.L3 = (int)d[2];")
        w("    D.PL = D.L1 * D.L2 * D.L3;")
        w(f"    D.NS = {self.b} * D.PL;")
        w("    A = G;")
        w(f"    A.spins = ({T} *)G.spins + d[6];")
        w("    A.mask = G.mask + d[7];")
        w("    A.k1 = G.k1 + d[7];")
        w("    A.site_ku = G.site_ku + d[7];")
        w(f"    A.site_axis = G.site_axis + d[7] * {n};")
        w("    A.beta = G.beta + d[4];")
        if self.proposal == "cone":
            w("    A.width = G.width + d[4];")
        if n == 1:
            w(f"    A.table = (const float *)G.table + d[4] * {total};")
        w("}")
        w()
        w("DEV uint32_t bkey(int64_t seed, int64_t sweep) {")
        w("    return (uint32_t)((seed * 1000003LL + sweep * 7919LL + 12345LL) % 2147483629LL);")
        w("}")
        w()
        w("DEV void bcoord(uint32_t c, int &x, int &y, int &z) {")
        w(f"    x = (int)(c & {(1 << COORD_BITS) - 1}u); y = (int)((c >> {COORD_BITS}) & {(1 << COORD_BITS) - 1}u); "
          f"z = (int)(c >> {2 * COORD_BITS});")
        w("}")
        w()

    def emit_class_pass(self, k: int, a: int, lo: str, hi: str, step: str, ind: str) -> None:
        """Update the sites list[lo .. hi) of class k (sublattice a), one per lane, `step` apart."""
        w = self.w
        w(f"{ind}for (int i = {lo}; i < {hi}; i += {step}) {{")
        w(f"{ind}    int x, y, z; bcoord(list[i], x, y, z);")
        w(f"{ind}    const int own = {self._mul(a, 'D.PL')} + (z * D.L2 + y) * D.L1 + x;")
        w(f"{ind}    upd_{a}(A, S, S, rbase, rep, own, x, y, z, key, D);")
        w(f"{ind}}}")

    def emit_chain_prologue(self, idx: str) -> None:
        w, n, T = self.w, self.n, self.T
        w(f"    const int32_t *tl = G.tiles + 4 * (int64_t)({idx});")
        w("    const int sys = tl[0], rep = tl[1];")
        w("    const int64_t *d = bdesc(G, sys);")
        w("    bdims D; sck_args A; bsys(G, d, D, A);")
        w(f"    {T} *S = ({T} *)A.spins;")
        w(f"    const int64_t rbase = (int64_t)rep * {n} * D.NS;")
        w("    const uint32_t *list = blist(G);")

    def generate_batch(self) -> BatchModule:
        s, w = self.s, self.w
        w(HEADER)
        w("#pragma clang fp contract(off)")
        w(f"// batch of systems sharing one structure: n={s.n} crystal {s.crystal.name} b={s.b} "
          f"periodic={s.periodic}; classes {self.classes}; mode {self.mode}; proposal {self.proposal}")
        w()
        self.emit_batch_helpers()
        for a in range(self.b):
            self.emit_update_plain(a)
            w()
        # class kernels: persistent workgroups stride over the tiles (system, replica, list range) of one class
        for k, (c, a) in enumerate(self.classes):
            w(f'extern "C" __global__ void __launch_bounds__({CLASS_BOUND}) sck_bclass_{k}(sck_args G) {{')
            w("    const int64_t sweep = G.state[1] + G.sweep_offset;")
            w("    for (int t = blockIdx.x; t < G.ntile; t += gridDim.x) {")
            w("        const int32_t *tl = G.tiles + 4 * (int64_t)t;")
            w("        const int sys = tl[0], rep = tl[1], lo = tl[2], hi = tl[3];")
            w("        const int64_t *d = bdesc(G, sys);")
            w("        bdims D; sck_args A; bsys(G, d, D, A);")
            w("        const uint32_t key = bkey(d[5], sweep);")
            w(f"        {self.T} *S = ({self.T} *)A.spins;")
            w(f"        const int64_t rbase = (int64_t)rep * {self.n} * D.NS;")
            w("        const uint32_t *list = blist(G);")
            self.emit_class_pass(k, a, "lo + (int)threadIdx.x", "hi", "(int)blockDim.x", "        ")
            w("    }")
            w("}")
            w()
        # chain kernels: one workgroup runs all classes of A.nsweep sweeps of one (system, replica) chain;
        # chains are independent, so barriers inside the workgroup replace the kernel boundaries
        for nt in CHAIN_BLOCKS:
            w(f'extern "C" __global__ void __launch_bounds__({nt}) sck_bchain_{nt}(sck_args G) {{')
            w("    if ((int)blockIdx.x >= G.ntile) return;")
            self.emit_chain_prologue("blockIdx.x")
            w("    if (tl[3] < 0) return;                    // an empty slot of the work list")
            w("    for (int sw = 0; sw < G.nsweep; ++sw) {")
            w("        const uint32_t key = bkey(d[5], G.state[1] + G.sweep_offset + sw);")
            for k, (c, a) in enumerate(self.classes):
                w("        {")
                w(f"            const int lo = (int)d[{8 + 2 * k}], hi = lo + (int)d[{9 + 2 * k}];")
                self.emit_class_pass(k, a, "lo + (int)threadIdx.x", "hi", "(int)blockDim.x", "            ")
                w("        }")
                w("        __syncthreads();")
            w("    }")
            w("}")
            w()
        # sequential chain: one thread per (system, replica), classes in order, sites in raster order
        w('extern "C" __global__ void __launch_bounds__(64) sck_bseq(sck_args G) {')
        w("    const int g = blockIdx.x * blockDim.x + threadIdx.x;")
        w("    if (g >= G.ntile) return;")
        self.emit_chain_prologue("g")
        w("    const uint32_t key = bkey(d[5], G.state[1] + G.sweep_offset);")
        for k, (c, a) in enumerate(self.classes):
            w("    {")
            w(f"        const int lo = (int)d[{8 + 2 * k}], hi = lo + (int)d[{9 + 2 * k}];")
            self.emit_class_pass(k, a, "lo", "hi", "1", "        ")
            w("    }")
        w("}")
        w()
        self.emit_batch_observe()
        w('extern "C" __global__ void sck_bump(sck_args A) {')
        w("    if (threadIdx.x == 0 && blockIdx.x == 0) ((int64_t *)A.state)[1] += A.sweep_offset;")
        w("}")
        src = "\n".join(self.lines) + "\n"
        kernels = [f"sck_bclass_{k}" for k in range(len(self.classes))] + \
                  [f"sck_bchain_{nt}" for nt in CHAIN_BLOCKS] + ["sck_bseq", "sck_bobs", "sck_bump"]
        _, _, total = ising_layout(s) if s.n == 1 else ([], [], 0)
        return BatchModule(src, kernels, self.classes, table_entries=total,
                           key=hashlib.sha256(src.encode()).hexdigest()[:16])

    def emit_batch_observe(self) -> None:
        """sck_bobs: float64 energy and magnetization of each (system, replica), one workgroup each (tiles in
        global replica order), written to obs[g] and obs[R + g * b n ...] like the single-system kernels."""
        w, n, b = self.w, self.n, self.b
        NV = 1 + b * n
        w(f'extern "C" __global__ void __launch_bounds__({BLOCK}) sck_bobs(sck_args G) {{')
        w("    const int g = blockIdx.x;")
        w("    if (g >= G.ntile) return;")
        self.emit_chain_prologue("g")
        w(f"    double acc[{NV}];")
        w(f"    for (int k = 0; k < {NV}; ++k) acc[k] = 0.0;")
        w("    for (int site = threadIdx.x; site < D.NS; site += blockDim.x) {")
        self.emit_site_energy(lambda k, q: f"(double)S[rbase + (int64_t){self._mul(k, 'D.NS')} + {q}]")
        w("    }")
        w(f"    __shared__ double red[{BLOCK // 64}][{NV}];")
        w(f"    for (int k = 0; k < {NV}; ++k) {{")
        w("        double v = acc[k];")
        w("        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);")
        w("        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = v;")
        w("    }")
        w("    __syncthreads();")
        w(f"    if (threadIdx.x < {NV}) {{")
        w("        double v = 0.0;")
        w(f"        for (int i = 0; i < {BLOCK // 64}; ++i) v += red[i][threadIdx.x];")
        w("        if (threadIdx.x == 0) G.obs[g] = v;")
        w(f"        else G.obs[G.nrep + (int64_t)g * {NV - 1} + (threadIdx.x - 1)] = v;")
        w("    }")
        w("}")
        w()


def generate_batch(systems, mode: int, proposal: str = "uniform", learned_meta=None, wide: bool = False) -> BatchModule:
    """HIP source of the batch kernels for systems sharing one Hamiltonian structure (checked by the caller)."""
    if mode not in (MODE_INT8, MODE_F32):
        raise ValueError("batched systems run unpacked int8 (Ising) or float32 spins")
    if proposal not in ("uniform", "cone", "learned"):
        raise ValueError("proposal must be 'uniform', 'cone' or 'learned'")
    if proposal != "uniform" and systems[0].n == 1:
        raise ValueError("a cone proposal needs a vector model (n = 2 or 3)")
    return _BatchGen([as_system(x) for x in systems], mode, proposal, learned_meta, wide).generate_batch()


def fits_resident(sys_: System, mode: int) -> bool:
    """Whether one replica's lattice fits the LDS-resident kernel."""
    return _Gen(as_system(sys_), mode, 1, "uniform").resident_plan() is not None


def generate(sys_: System, mode: int, batch: int, proposal: str = "uniform", learned_meta=None,
             replica_major: bool = False, rm_tile=None, rm_chunks=None, rm_sites=None,
             resident_threads=None, rm_block=None, tiled: bool = False, tile_dims=None) -> Module:
    """HIP source and launch plans of one system (the counterpart of SCK/generator.py:150 `source` and
    SCK/sampler.py:191-232 `_launch`)."""
    if proposal not in ("uniform", "cone", "learned"):
        raise ValueError("proposal must be 'uniform', 'cone' or 'learned'")
    if proposal != "uniform" and sys_.n == 1:
        raise ValueError("a cone proposal needs a vector model (n = 2 or 3)")
    return _Gen(as_system(sys_), mode, batch, proposal, learned_meta, replica_major, rm_tile, rm_chunks, rm_sites,
                resident_threads, rm_block, tiled, tile_dims).generate()


def fits_tiled(sys_: System, mode: int, tile_dims=None) -> bool:
    """Whether the tiled sweep kernel applies to the system (see _Gen.tiled_plan)."""
    return _Gen(as_system(sys_), mode, 1, "uniform", tile_dims=tile_dims).tiled_plan() is not None
