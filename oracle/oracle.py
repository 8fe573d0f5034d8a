"""ctypes front of liboracle.so and numpy restatements used by the checks (test infrastructure only).

Every function cites the reference lines it restates. Spins cross this boundary in the reference layout
[B, n, b, L3, L2, L1].
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

__all__ = ["lib", "build", "OracleSystem", "chain", "sync", "bitwise_chain", "observe", "colouring_ok",
           "philox", "u01", "sweep_seed", "p_table", "p_fixed", "init_spins", "cone", "learned", "pack", "unpack",
           "numpy_exp32", "PROP_UNIFORM", "PROP_CONE"]

PROP_UNIFORM, PROP_CONE = 0, 1
_lib = None


def build() -> Path:
    """Compile liboracle.so from oracle.c (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists() or LIB.stat().st_mtime < (HERE / "oracle.c").stat().st_mtime:
            build()
        _lib = C.CDLL(str(LIB))
        _lib.oracle_chain.restype = C.c_int64
        _lib.oracle_sweep_seed.restype = C.c_int64
        _lib.oracle_u01.restype = C.c_float
        _lib.oracle_colouring_ok.restype = C.c_int
    return _lib


class _Sys(C.Structure):
    _fields_ = [("n", C.c_int32), ("b", C.c_int32), ("D", C.c_int32), ("L", C.c_int32 * 3),
                ("periodic", C.c_int32 * 3), ("nbond", C.c_int32), ("bond", C.c_void_p), ("bondJ", C.c_void_p),
                ("has_extra", C.c_void_p), ("extra", C.c_void_p), ("has_h", C.c_int32), ("h", C.c_double * 3),
                ("mask", C.c_void_p), ("has_cubic", C.c_int32), ("k1", C.c_void_p), ("k2", C.c_double),
                ("has_uniaxial", C.c_int32), ("du", C.c_double), ("du_axis", C.c_double * 3),
                ("has_site", C.c_int32), ("site_ku", C.c_void_p), ("site_axis", C.c_void_p),
                ("barker", C.c_int32), ("r", C.c_int32), ("m", C.c_int32 * 4), ("K", (C.c_int32 * 3) * 4),
                ("o", (C.c_int32 * 4) * 8)]


class _Prop(C.Structure):
    _fields_ = [("kind", C.c_int32), ("width", C.c_void_p), ("width_table", C.c_void_p), ("n", C.c_int32 * 3),
                ("s", C.c_float * 3), ("axis", C.c_float * 3)]


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class OracleSystem:
    """The oracle's descriptor of a System; keeps the numpy buffers it points to alive."""

    def __init__(self, s):
        self.s = s
        cr = s.crystal
        D = cr.D
        self.keep = []
        bond = np.zeros((max(1, len(cr.bonds)), 6), np.int32)
        bondJ = np.zeros(max(1, len(cr.bonds)), np.float64)
        has_extra = np.zeros(max(1, len(cr.bonds)), np.int32)
        extra = np.zeros((max(1, len(cr.bonds)), 3, 3), np.float64)
        for i, bd in enumerate(cr.bonds):
            d = list(bd.delta) + [0] * (3 - D)
            bond[i] = (bd.alpha, bd.alpha2, d[0], d[1], d[2], bd.shell)
            bondJ[i] = float(s.J.get(bd.shell, 0.0))
            if i in s.extra:
                has_extra[i] = 1
                E = np.asarray(s.extra[i], np.float64)
                extra[i, :E.shape[0], :E.shape[1]] = E
        mask = np.ascontiguousarray(s.site_mask().astype(np.int8).reshape(-1))
        k1 = np.ascontiguousarray((s.cubic_k1 if s.cubic_k1 is not None else np.zeros(s.shape)).astype(np.float32).reshape(-1))
        ku = np.ascontiguousarray((s.site_ku if s.site_ku is not None else np.zeros(s.shape)).astype(np.float32).reshape(-1))
        ax = np.ascontiguousarray((s.site_axis if s.site_axis is not None else np.zeros(s.shape + (s.n,))).astype(np.float32).reshape(-1))
        self.keep += [bond, bondJ, has_extra, extra, mask, k1, ku, ax]
        st = _Sys()
        st.n, st.b, st.D = s.n, s.b, D
        for k, v in enumerate((s.L1, s.L2, s.L3)):
            st.L[k] = v
        for k in range(3):
            st.periodic[k] = int(s.periodic[k]) if k < D else 0
        st.nbond = len(cr.bonds)
        st.bond, st.bondJ, st.has_extra, st.extra = _ptr(bond), _ptr(bondJ), _ptr(has_extra), _ptr(extra)
        st.has_h = int(s.h is not None)
        if s.h is not None:
            for k, v in enumerate(s.h):
                st.h[k] = float(v)
        st.mask = _ptr(mask)
        st.has_cubic = int(s.cubic_k1 is not None)
        st.k1, st.k2 = _ptr(k1), float(s.cubic_k2)
        st.has_uniaxial = int(s.uniaxial is not None)
        if s.uniaxial is not None:
            st.du = float(s.uniaxial[0])
            for k, v in enumerate(s.uniaxial[1]):
                st.du_axis[k] = float(v)
        st.has_site = int(s.site_ku is not None)
        st.site_ku, st.site_axis = _ptr(ku), _ptr(ax)
        st.barker = int(s.acceptance == "barker")
        col = s.colouring
        st.r = len(col.m)
        for r, mr in enumerate(col.m):
            st.m[r] = mr
            for k in range(D):
                st.K[r][k] = col.K[r][k]
        for a in range(s.b):
            for r in range(len(col.m)):
                st.o[a][r] = col.o[a][r]
        self.st = st
        self.nsite = s.b * s.L3 * s.L2 * s.L1


def numpy_exp32(x: float) -> float:
    """float32 exp as the reference's CPU chain evaluates tl.exp (numpy on float32)."""
    with np.errstate(over="ignore"):
        return float(np.exp(np.asarray([x], np.float32))[0])


_EXP_CB = C.CFUNCTYPE(C.c_float, C.c_float)(numpy_exp32)


def _prop(kind: int = PROP_UNIFORM, width=None, table=None, n=(1, 1, 1), s=(0.0, 0.0, 0.0), axis=(0.0, 0.0, 1.0)):
    p = _Prop()
    p.kind = kind
    keep = []
    if width is not None:
        w = np.ascontiguousarray(np.asarray(width, np.float32).reshape(-1))
        keep.append(w)
        p.width = _ptr(w)
    if table is not None:
        t = np.ascontiguousarray(np.asarray(table, np.float32).reshape(-1))
        keep.append(t)
        p.width_table = _ptr(t)
        for d in range(3):
            p.n[d], p.s[d] = int(n[d]), float(s[d])
    for k, v in enumerate(axis):
        p.axis[k] = float(v)
    return p, keep


def chain(osys: OracleSystem, spins: np.ndarray, beta: np.ndarray, base_seed: int, sweep0: int, nsweeps: int,
          reps=None, prop=None, class_range=(0, -1), margins: bool = False, threads: int = 1):
    """Sequential chain (SCK/sampler.py:281-293 with the update of SCK/generator.py:150-319).

    spins: (nrep, n, b, L3, L2, L1) for the replicas `reps` (default 0..nrep-1); returned updated as float32.
    beta: per replica of the whole batch (indexed by replica number). Returns (spins, accepted[, margins])."""
    s = osys.s
    S = np.ascontiguousarray(np.asarray(spins, np.float32).reshape(spins.shape[0], s.n, -1)).copy()
    nrep = S.shape[0]
    reps = np.ascontiguousarray(np.arange(nrep) if reps is None else np.asarray(reps), np.int64)
    beta = np.ascontiguousarray(np.asarray(beta, np.float32).reshape(-1))
    p, keep = prop if prop is not None else _prop()
    mg = np.zeros((nrep, osys.nsite), np.float64) if margins else None
    acc = lib().oracle_chain(C.byref(osys.st), C.c_void_p(_ptr(S)), C.c_void_p(_ptr(reps)), C.c_int64(nrep),
                             C.c_void_p(_ptr(beta)), C.c_int64(int(base_seed)), C.c_int64(int(sweep0)),
                             C.c_int64(int(nsweeps)), C.byref(p), _EXP_CB, C.c_int(class_range[0]),
                             C.c_int(class_range[1]), C.c_void_p(_ptr(mg)) if margins else None, C.c_int(threads))
    out = S.reshape(spins.shape)
    return (out, acc, mg) if margins else (out, acc)


def sync(osys: OracleSystem, spins: np.ndarray, beta: np.ndarray, base_seed: int, sweep0: int, nsweeps: int,
         reps=None, prop=None):
    """Synchronous control (SCK/sampler.py:271-279)."""
    s = osys.s
    S = np.ascontiguousarray(np.asarray(spins, np.float32).reshape(spins.shape[0], s.n, -1)).copy()
    reps = np.ascontiguousarray(np.arange(S.shape[0]) if reps is None else np.asarray(reps), np.int64)
    beta = np.ascontiguousarray(np.asarray(beta, np.float32).reshape(-1))
    p, keep = prop if prop is not None else _prop()
    lib().oracle_sync(C.byref(osys.st), C.c_void_p(_ptr(S)), C.c_void_p(_ptr(reps)), C.c_int64(S.shape[0]),
                      C.c_void_p(_ptr(beta)), C.c_int64(int(base_seed)), C.c_int64(int(sweep0)),
                      C.c_int64(int(nsweeps)), C.byref(p), _EXP_CB)
    return S.reshape(spins.shape)


def cone(width) -> tuple:
    """Cone-proposal descriptor with one half-angle per replica."""
    return _prop(PROP_CONE, width=width)


def learned(table, n, s, axis) -> tuple:
    """Learned-width descriptor: width = table[i0, i1, i2] over quantised features (oracle.c, cone_width)."""
    return _prop(PROP_CONE, table=table, n=n, s=s, axis=axis)


# ------------------------------------------------------------------------------------------ bitwise
def _bitwise_params(s):
    """(J, h, Z) of a system the bitwise kernel accepts (SCK/bitwise.py:59-76)."""
    Js = {float(s.J.get(bd.shell, 0.0)) for bd in s.crystal.bonds} - {0.0}
    h = float(s.h[0]) if s.h is not None else 0.0
    Z = max(sum(1 for bd in s.crystal.bonds if bd.alpha == a and s.J.get(bd.shell, 0.0) != 0.0) for a in range(s.b))
    return Js.pop(), h, Z


def _p64(s, beta) -> np.ndarray:
    """p[word, s, z_eff, k] in float64 (SCK/bitwise.py:96-110): dE = 2J(z_eff - 2k) + 2h sigma."""
    J, h, Z = _bitwise_params(s)
    beta = np.asarray(beta, np.float64).reshape(-1)
    p = np.zeros((beta.size, 2, Z + 1, Z + 1), np.float64)
    for si in (0, 1):
        for ze in range(Z + 1):
            k = np.arange(ze + 1)
            x = beta[:, None] * (2.0 * J * (ze - 2 * k) + 2.0 * h * (2 * si - 1))[None, :]
            p[:, si, ze, :ze + 1] = (0.5 * (1.0 - np.tanh(0.5 * x)) if s.acceptance == "barker"
                                     else np.exp(-np.maximum(x, 0.0)))
    return p


def p_table(s, beta) -> np.ndarray:
    """Shared-mode table, float32 (SCK/bitwise.py:79-84)."""
    return _p64(s, beta).astype(np.float32)


def p_fixed(s, beta) -> np.ndarray:
    """Independent-mode table P = floor(p 2^32) clipped to [0, 2^32], int64 (SCK/bitwise.py:87-93)."""
    return np.clip(np.floor(_p64(s, beta) * 2.0 ** 32).astype(np.int64), 0, 2 ** 32)


def bitwise_chain(osys: OracleSystem, spins: np.ndarray, ptab: np.ndarray, base_seed: int, sweep0: int,
                  nsweeps: int, independent: bool = False, threads: int = 1, word_offset: int = 0) -> np.ndarray:
    """Sequential bitwise chain on unpacked spins (B, b, L3, L2, L1) int8 (SCK/bitwise.py:194-246). The spins and
    table may cover words word_offset .. word_offset + ceil(B/32) - 1 of a larger batch."""
    s = osys.s
    J, h, Z = _bitwise_params(s)
    S = np.ascontiguousarray(np.asarray(spins, np.int8).reshape(spins.shape[0], -1)).copy()
    B = S.shape[0]
    if independent:
        pi = np.ascontiguousarray(np.asarray(ptab, np.int64))
        pf = np.zeros(1, np.float32)
    else:
        pf = np.ascontiguousarray(np.asarray(ptab, np.float32))
        pi = np.zeros(1, np.int64)
    lib().oracle_bitwise(C.byref(osys.st), C.c_void_p(_ptr(S)), C.c_int64(B), C.c_void_p(_ptr(pf)),
                         C.c_void_p(_ptr(pi)), C.c_int(Z + 1), C.c_int(int(h != 0.0)), C.c_int(int(independent)),
                         C.c_int64(int(base_seed)), C.c_int64(int(sweep0)), C.c_int64(int(nsweeps)), C.c_int(threads),
                         C.c_int64(int(word_offset)))
    return S.reshape(spins.shape)


# -------------------------------------------------------------------------------------- observables
def observe(osys: OracleSystem, spins: np.ndarray):
    """(E (B,), M (B, b, n)) in float64 (SCK/sampler.py:296-345)."""
    s = osys.s
    S = np.ascontiguousarray(np.asarray(spins, np.float64).reshape(spins.shape[0], s.n, -1))
    B = S.shape[0]
    E = np.zeros(B, np.float64)
    M = np.zeros((B, s.b, s.n), np.float64)
    lib().oracle_observe(C.byref(osys.st), C.c_void_p(_ptr(S)), C.c_int64(B), C.c_void_p(_ptr(E)), C.c_void_p(_ptr(M)))
    return E, M


def colouring_ok(osys: OracleSystem) -> bool:
    return bool(lib().oracle_colouring_ok(C.byref(osys.st)))


# ------------------------------------------------------------------------------------------- random
def philox(key, counters, rounds: int = 10) -> np.ndarray:
    """Raw Philox-4x32 outputs (n, 4) for counters (n, 4) and key (k0, k1)."""
    c = np.ascontiguousarray(np.asarray(counters, np.uint32).reshape(-1, 4))
    cols = [np.ascontiguousarray(c[:, j]) for j in range(4)]
    out = np.zeros((c.shape[0], 4), np.uint32)
    lib().oracle_philox(C.c_uint32(key[0]), C.c_uint32(key[1]), *[C.c_void_p(_ptr(x)) for x in cols],
                        C.c_int64(c.shape[0]), C.c_int(rounds), C.c_void_p(_ptr(out)))
    return out


def u01(x: int) -> float:
    return float(lib().oracle_u01(C.c_uint32(x)))


def sweep_seed(base: int, sweep: int) -> int:
    return int(lib().oracle_sweep_seed(C.c_int64(base), C.c_int64(sweep)))


# ----------------------------------------------------------------------------------------- storage
def init_spins(s, B: int, kind: str = "random", seed: int = 0) -> np.ndarray:
    """Initial spins (B, n, b, L3, L2, L1) float32 exactly as the reference's Sampler.init_spins draws them
    with the torch CPU generator (SCK/sampler.py:152-167), masked sites zero."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    shape = (B, s.n) + s.shape
    if kind == "ordered":
        x = torch.zeros(shape)
        x[:, 0] = 1.0
    elif s.n == 1:
        x = torch.randint(0, 2, shape, generator=g).float() * 2 - 1
    else:
        x = torch.randn(shape, generator=g)
        x = x / x.norm(dim=1, keepdim=True)
    return (x * torch.as_tensor(s.site_mask(), dtype=torch.float32)).numpy()


def pack(x: np.ndarray) -> np.ndarray:
    """(B, 1, b, L3, L2, L1) +-1 -> (W, 1, b, L3, L2, L1) int32 words, bit j of word w = replica 32w + j
    (1 = +1), last word zero-padded (SCK/sampler.py:169-178)."""
    B = x.shape[0]
    W = -(-B // 32)
    bits = np.zeros((W * 32,) + x.shape[1:], np.uint64)
    bits[:B] = (np.asarray(x) > 0).astype(np.uint64)
    bits = bits.reshape((W, 32) + x.shape[1:])
    sh = np.arange(32, dtype=np.uint64).reshape((1, 32) + (1,) * (x.ndim - 1))
    return (bits << sh).sum(axis=1).astype(np.uint32).view(np.int32)


def unpack(words: np.ndarray, B: int, mask: np.ndarray) -> np.ndarray:
    """Inverse of pack, int8, masked sites 0 (SCK/sampler.py:180-188)."""
    w = np.asarray(words).view(np.uint32).astype(np.uint64)
    sh = np.arange(32, dtype=np.uint64).reshape((1, 32) + (1,) * (w.ndim - 1))
    bits = (w[:, None] >> sh) & np.uint64(1)
    x = (bits.astype(np.int64) * 2 - 1).reshape((-1,) + w.shape[1:])[:B]
    return (x * np.asarray(mask, np.int64)).astype(np.int8)

