"""ctypes binding of libsck.so (include/sck.h) and its in-tree build.

torch is imported before the library is loaded so that libsck's libamdhip64.so.7 / libhiprtc.so.7 resolve
to the copies torch already mapped: one HIP runtime per process, so torch's streams and pointers are valid
handles for libsck.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import torch  # noqa: F401  (must precede loading libsck: shared HIP runtime)

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent                      # spin-colour-kernel_amd/
REPO = ROOT.parent
INCLUDE = REPO / "include"
LIB = PKG / "libsck.so"
SOURCES = [ROOT / "csrc" / "sck.cpp", ROOT / "csrc" / "mlp.hip"]
CACHE = Path(os.environ.get("SCK_AMD_CACHE", ROOT / "_kcache"))

SCK_ABI_VERSION = 6
OBS_SCRATCH_BYTES = 32 << 20
CAPTURE_OBSERVE = 1        # sck_graph_capture flags (include/sck.h)
CAPTURE_IN_DST = 2          # SCK_OBS_SCRATCH_BYTES of include/sck.h


class Args(C.Structure):
    _fields_ = [("spins", C.c_void_p), ("dst", C.c_void_p), ("mask", C.c_void_p), ("k1", C.c_void_p),
                ("beta", C.c_void_p), ("table", C.c_void_p), ("state", C.c_void_p), ("width", C.c_void_p),
                ("site_ku", C.c_void_p), ("site_axis", C.c_void_p), ("obs", C.c_void_p),
                ("obs_scratch", C.c_void_p), ("sweep_offset", C.c_int64), ("nrep", C.c_int32),
                ("nsweep", C.c_int32), ("aux", C.c_void_p), ("tiles", C.c_void_p), ("ntile", C.c_int32),
                ("rep_base", C.c_int32), ("track", C.c_void_p), ("track_parts", C.c_void_p)]


class LaunchC(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("grid", C.c_uint32 * 3), ("block", C.c_uint32 * 3), ("shmem", C.c_uint32),
                ("tiles", C.c_void_p), ("ntile", C.c_int32), ("rep_base", C.c_int32)]


class Desc(C.Structure):
    _fields_ = [("abi_version", C.c_int32), ("mode", C.c_int32), ("n", C.c_int32), ("b", C.c_int32),
                ("L1", C.c_int32), ("L2", C.c_int32), ("L3", C.c_int32), ("batch", C.c_int32),
                ("words", C.c_int32), ("obs_chunks", C.c_int32),
                ("kernel_names", C.POINTER(C.c_char_p)), ("n_kernels", C.c_int32),
                ("sweep", C.POINTER(LaunchC)), ("n_sweep", C.c_int32),
                ("ref", C.POINTER(LaunchC)), ("n_ref", C.c_int32),
                ("sync", C.POINTER(LaunchC)), ("n_sync", C.c_int32),
                ("obs", C.POINTER(LaunchC)), ("n_obs", C.c_int32),
                ("bump", C.POINTER(LaunchC)), ("resident", C.POINTER(LaunchC)), ("rep_blocks", C.c_int32),
                ("obs_blocks", C.c_int32), ("tiled", C.POINTER(LaunchC)), ("track_fold", C.POINTER(LaunchC))]


#: every symbol include/sck.h declares
EXPORTS = ["sck_version", "sck_last_error", "sck_create", "sck_precompile", "sck_destroy", "sck_sweep",
           "sck_sweep_timed", "sck_reference_sweep", "sck_sync_sweep", "sck_observe", "sck_observe_noalloc",
           "sck_graph_capture", "sck_graph_replay", "sck_graph_destroy", "sck_time_replays", "sck_mlp_nparam",
           "sck_mlp_workspace_bytes", "sck_mlp_train_step", "sck_mlp_table", "sck_sample_sites"]


class MlpSample(C.Structure):
    _fields_ = [("s", C.c_float * 3), ("v", C.c_float * 3), ("beta", C.c_float), ("k1", C.c_float),
                ("ku", C.c_float), ("axis", C.c_float * 3)]


class MlpConfig(C.Structure):
    _fields_ = [("n", C.c_int32), ("ndraw", C.c_int32), ("seed", C.c_int32), ("wmin", C.c_float),
                ("wmax", C.c_float), ("k2", C.c_float), ("du", C.c_float), ("axis", C.c_float * 3),
                ("xs", C.c_float * 3)]


class AdamConfig(C.Structure):
    _fields_ = [("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("step", C.c_int32)]


class WidthGrid(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("s", C.c_float * 3)]


def build(verbose: bool = False) -> Path:
    """Compile libsck.so in-tree with hipcc for gfx950: the C-ABI launcher and the static kernels of the
    learned-proposal training (the sweep kernels are generated per system and compiled by hipRTC)."""
    cmd = ["hipcc", "-shared", "-fPIC", "-O3", "-std=c++17", "--offload-arch=gfx950", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", f"-I{INCLUDE}", "-x", "hip",
           *map(str, SOURCES), "-o", str(LIB) + ".tmp", "-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building libsck.so failed:\n{r.stdout}\n{r.stderr}")
    os.replace(str(LIB) + ".tmp", LIB)
    if verbose:
        print(" ".join(cmd))
    return LIB


_lib = None


def lib():
    """The loaded libsck.so; raises when it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(str(LIB))
        L.sck_last_error.restype = C.c_char_p
        L.sck_version.restype = C.c_int
        for name in EXPORTS:
            getattr(L, name)
        L.sck_create.argtypes = [C.POINTER(Desc), C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]
        L.sck_precompile.argtypes = [C.c_char_p, C.c_char_p]
        L.sck_destroy.argtypes = [C.c_void_p]
        L.sck_sweep.argtypes = [C.c_void_p, C.POINTER(Args), C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
        L.sck_sweep_timed.argtypes = [C.c_void_p, C.POINTER(Args), C.c_int32, C.c_void_p, C.POINTER(C.c_float),
                                      C.POINTER(C.c_int32)]
        L.sck_reference_sweep.argtypes = [C.c_void_p, C.POINTER(Args), C.c_int32, C.c_void_p]
        L.sck_sync_sweep.argtypes = [C.c_void_p, C.POINTER(Args), C.c_int32, C.c_int64, C.c_void_p]
        L.sck_observe.argtypes = [C.c_void_p, C.POINTER(Args), C.c_void_p]
        L.sck_observe_noalloc.argtypes = [C.c_void_p, C.POINTER(Args), C.c_void_p]
        L.sck_graph_capture.argtypes = [C.c_void_p, C.POINTER(Args), C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]
        L.sck_graph_replay.argtypes = [C.c_void_p, C.c_void_p]
        L.sck_graph_destroy.argtypes = [C.c_void_p]
        L.sck_time_replays.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_float)]
        L.sck_mlp_nparam.restype = C.c_int
        L.sck_mlp_workspace_bytes.restype = C.c_int64
        L.sck_mlp_workspace_bytes.argtypes = [C.c_int32, C.c_int32]
        L.sck_mlp_train_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                         C.POINTER(MlpConfig), C.POINTER(AdamConfig), C.c_void_p, C.c_void_p,
                                         C.c_void_p]
        L.sck_mlp_table.argtypes = [C.c_void_p, C.POINTER(MlpConfig), C.POINTER(WidthGrid), C.c_void_p, C.c_void_p]
        L.sck_sample_sites.argtypes = [C.c_void_p, C.POINTER(Args), C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        if L.sck_version() != SCK_ABI_VERSION:
            raise RuntimeError("libsck.so ABI version mismatch: rebuild it")
        _lib = L
    return _lib


def check(rc: int, what: str = "libsck") -> None:
    if rc != 0:
        msg = lib().sck_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(f"{what}: {msg}")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def launches(ls) -> tuple:
    arr = (LaunchC * max(1, len(ls)))()
    for i, l in enumerate(ls):
        arr[i].kernel = l.kernel
        for k in range(3):
            arr[i].grid[k] = int(l.grid[k])
            arr[i].block[k] = int(l.block[k])
        arr[i].shmem = int(l.shmem)
        arr[i].tiles = getattr(l, "tiles_ptr", None)
        arr[i].ntile = int(getattr(l, "ntile", 0))
        arr[i].rep_base = int(getattr(l, "rep_base", 0))
    return arr, len(ls)


def precompile(src: str, cache_dir: Path | None = None) -> None:
    """Compile one generated module into the code-object cache (no GPU needed)."""
    check(lib().sck_precompile(src.encode(), str(cache_dir or CACHE).encode()), "hipRTC precompile")
