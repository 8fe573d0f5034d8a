"""Host-side system model, generator and C-ABI library (CPU only: nothing here launches a kernel).

Known answers follow the reference's tests (tests/test_core.py:12-36, 118-130; tests/test_bitwise.py:119-147).
"""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from sck_amd import colouring as co
from sck_amd import crystals as cr
from sck_amd import hipgen as hg
from sck_amd import system as sy

REPO = Path(__file__).resolve().parents[1]

EXPECTED = {("square", (1,)): (2, 2), ("sc", (1,)): (2, 2), ("sc", (1, 2)): (4, 4), ("sc", (1, 2, 3, 4)): (9, 8),
            ("bcc", (1,)): (2, 2), ("fcc", (1,)): (4, 4), ("triangular", (1,)): (3, 3),
            ("honeycomb", (1,)): (2, 2), ("kagome", (1,)): (3, 3)}


@pytest.mark.parametrize("key", sorted(EXPECTED))
def test_colouring_counts_and_admissibility(key):
    c = cr.make(*key)
    col = co.search(c)
    assert co.admissible(col, c)
    assert (col.n_colours, co.clique_number(c)) == EXPECTED[key]


def test_bond_counts():
    assert [len(cr.make(n).bonds) for n in ("square", "sc", "bcc", "fcc", "triangular", "kagome")] == [4, 6, 16, 48, 6, 12]


def test_seam_condition_rejects_odd_box():
    c = cr.make("triangular")
    col = co.search(c)
    assert not co.seam_ok(col, (4, 4), (True, True))
    with pytest.raises(ValueError):
        sy.System("bad", 1, c, (4, 4), (True, True), {1: 1.0}, colouring=col)


def test_x_period_of_odd_periodic_box():
    s = sy.System("Ising sc odd periodic", 1, cr.make("sc"), (5, 5, 5), (True,) * 3, {1: 1.0})
    assert s.colouring.n_colours == 5 and hg.x_period(s.colouring) == 5


def test_large_boxes_colour_quickly_and_match_the_survey():
    """Triangular 2048^2 cannot take Z3 (2048 mod 3 != 0): four colours Z2 x Z2; bcc shells 1-2: four classes."""
    t = sy.xy("triangular", 2048)
    assert t.colouring.m == (2, 2) and len(t.classes()) == 4
    b = sy.System("bcc J1J2", 3, cr.make("bcc", (1, 2)), (32, 32, 32), (True,) * 3, {1: 1.0, 2: -0.2})
    assert b.colouring.n_colours == 4 and len(b.classes()) == 4
    s = sy.ising("sc", 128)
    assert len(s.classes()) == 2


def test_vectorised_colour_map_equals_the_definition():
    s = sy.nanodot(Rd=3.0, L=3)
    cm = s.colour_map()
    for a in range(s.b):
        for z in range(s.L3):
            for y in range(s.L2):
                for x in range(s.L1):
                    assert cm[a, z, y, x] == s.colouring.colour((x, y, z), a)


def test_nanodot_defaults_and_shells():
    s = sy.nanodot(Rd=2.5, L=3)
    assert s.J == {1: 1.0} and s.colouring.n_colours == 2 and len(s.extra) == 4
    s2 = sy.nanodot(Rd=2.5, L=3, shells=(1, 2, 4), J_shells={2: 0.3, 4: -0.1})
    assert s2.J == {1: 1.0, 2: 0.3, 4: -0.1} and s2.colouring.n_colours > 2


def test_ising_table_dE_values_square():
    """Square lattice, J = 1, h = 0: dE = 2 s sum_j s_j over sums -4..4 (indices of the table digits)."""
    dE = hg.ising_dE_values(sy.ising("square", 4))
    g = np.repeat(np.arange(-4, 5), 2)
    s = np.tile([-1, 1], 9)
    assert np.array_equal(dE, (2 * s * g).astype(np.float32))


def test_ising_table_two_groups_in_reference_order():
    s = sy.System("J1-J2", 1, cr.make("sc", (1, 2)), (4, 4, 4), (True,) * 3, {1: 1.0, 2: 0.3}, h=(0.1,))
    layouts, offsets, total = hg.ising_layout(s)
    assert [d[1] for d in layouts[0]] == [13, 25, 2] and total == 13 * 25 * 2
    dE = hg.ising_dE_values(s)
    # entry (g1=6, g2=12, s=+1): v = (0 + 1*6) + fp32(0.3)*12 + fp32(0.1), dE = 2 v in fp32
    v = np.float32(np.float32(0) + np.float32(1.0) * np.float32(6))
    v = np.float32(v + np.float32(0.3) * np.float32(12))
    v = np.float32(v + np.float32(0.1))
    assert dE[(12 * 25 + 24) * 2 + 1] == np.float32(2) * v


def test_metropolis_table_uses_float32_exp_of_float32_product():
    s = sy.ising("square", 4)
    beta = np.float32(1) / np.array([2.0, 3.0], np.float32)
    t = hg.ising_table(s, beta)
    dE = hg.ising_dE_values(s)
    assert np.array_equal(t, np.exp(((-beta)[:, None] * dE[None, :]).astype(np.float32)))


def test_generator_modes_emit_all_kernels():
    for s, mode in [(sy.ising("square", 8), hg.MODE_INT8), (sy.xy("square", 4), hg.MODE_F32),
                    (sy.nanodot(Rd=2.0, L=2), hg.MODE_F32), (sy.ising("square", 4), hg.MODE_PACKED),
                    (sy.ising("sc", 4), hg.MODE_BITWISE), (sy.ising("sc", 4), hg.MODE_BITWISE_INDEP)]:
        m = hg.generate(s, mode, 64)
        for k in m.kernels:
            assert re.search(rf"__global__ void .*\b{k}\(sck_args A[,)]", m.src), k
        assert len(m.sweep) == len(s.classes())


def test_library_exports_every_declared_symbol():
    from sck_amd import _lib as L
    header = (REPO / "include" / "sck.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*(sck_\w+)\(", header, re.M))
    assert declared == set(L.EXPORTS)
    lib = ctypes.CDLL(str(L.LIB))
    for name in declared:
        assert hasattr(lib, name)
    hdr = (Path(__file__).resolve().parents[1] / "include" / "sck.h").read_text()
    version = int(re.search(r"#define SCK_ABI_VERSION (\d+)", hdr).group(1))
    assert L.lib().sck_version() == version == L.SCK_ABI_VERSION


def test_create_rejects_incomplete_descriptors_before_any_device_work():
    """sck_create checks the descriptor before it compiles or loads anything: a missing counter-bump plan, a
    plan count without its array, a replica-block count that does not divide the sweep launches, or per-block
    observation groups that do not match the replica blocks."""
    from sck_amd import _lib as L
    lib = L.lib()
    names = (ctypes.c_char_p * 1)(b"sck_class_0")
    plan, cnt = L.launches([hg.Launch(0, (1, 1, 1))])

    def desc(**kw):
        d = L.Desc()
        d.abi_version = L.SCK_ABI_VERSION
        d.kernel_names = ctypes.cast(names, ctypes.POINTER(ctypes.c_char_p))
        d.n_kernels = 1
        d.rep_blocks = 1
        d.sweep = ctypes.cast(plan, ctypes.POINTER(L.LaunchC))
        d.n_sweep = cnt
        d.bump = ctypes.cast(plan, ctypes.POINTER(L.LaunchC))
        for k, v in kw.items():
            setattr(d, k, v)
        return d

    h = ctypes.c_void_p()
    for d, msg in [(desc(bump=None), b"launch plans missing"), (desc(n_ref=2), b"launch plans missing"),
                   (desc(rep_blocks=2), b"rep_blocks"), (desc(abi_version=L.SCK_ABI_VERSION + 1), b"ABI"),
                   (desc(obs_blocks=2, obs=ctypes.cast(plan, ctypes.POINTER(L.LaunchC)), n_obs=2, batch=2), b"obs_blocks")]:
        assert lib.sck_create(ctypes.byref(d), b"// unused", None, ctypes.byref(h)) != 0
        assert msg in lib.sck_last_error()
        assert not h.value


def test_hiprtc_compiles_generated_modules_without_a_device(tmp_path):
    from sck_amd import _lib as L
    for s, mode in [(sy.heisenberg("bcc", 2), hg.MODE_F32), (sy.ising("sc", 4), hg.MODE_BITWISE_INDEP)]:
        L.precompile(hg.generate(s, mode, 64).src, tmp_path)
    assert len(list(tmp_path.glob("*.co"))) == 2


def test_bad_source_reports_the_compiler_log(tmp_path):
    from sck_amd import _lib as L
    with pytest.raises(RuntimeError, match="hipRTC"):
        L.precompile('extern "C" __global__ void k() { this is not code }', tmp_path)


def test_sampler_refuses_cpu_device():
    from sck_amd.sampler import Sampler
    with pytest.raises(ValueError, match="HIP device"):
        Sampler(sy.ising("square", 4), 2, device="cpu")


@pytest.mark.parametrize("case", ["bcc32-8192", "tri64-2048", "sq30-192-tile", "sc10-320-tile", "bcc32-8192-R1",
                                  "bcc32-8192-R2", "bcc32-8192-site-waves", "sq30-192-site-waves",
                                  "bcc4-256-block64", "bcc4-512-block128-site-waves"])
def test_replica_major_grid_covers_every_site_and_replica_group_once(case):
    """The launch plan's grid with the kernel's decode (XCD share -> (replica group, tiled slot)) visits every
    (group, class slot) exactly once; ragged tiles and the padded XCD shares fall out of range."""
    from sck_amd import hipgen as G
    mk = {"bcc32-8192": (lambda: sy.System("b", 3, cr.make("bcc", (1, 2)), (32, 32, 32), (True,) * 3,
                                            {1: 1.0, 2: -0.2}), 8192, None),
          "tri64-2048": (lambda: sy.xy("triangular", 64), 2048, None),
          "sq30-192-tile": (lambda: sy.xy("square", 30), 192, (7, 5, 1)),
          "sc10-320-tile": (lambda: sy.heisenberg("sc", 10), 320, (3, 4, 3)),
          "bcc32-8192-R1": (lambda: sy.System("b", 3, cr.make("bcc", (1, 2)), (32, 32, 32), (True,) * 3,
                                              {1: 1.0, 2: -0.2}), 8192, (4, 4, 4), 1),
          "bcc32-8192-R2": (lambda: sy.System("b", 3, cr.make("bcc", (1, 2)), (32, 32, 32), (True,) * 3,
                                              {1: 1.0, 2: -0.2}), 8192, (8, 4, 2), 2),
          "bcc32-8192-site-waves": (lambda: sy.System("b", 3, cr.make("bcc", (1, 2)), (32, 32, 32), (True,) * 3,
                                                      {1: 1.0, 2: -0.2}), 8192, (4, 4, 4), 2, 4),
          "sq30-192-site-waves": (lambda: sy.xy("square", 30), 192, (7, 5, 1), 1, 4),
          "bcc4-256-block64": (lambda: sy.heisenberg("bcc", 4), 256, None, None, 1, 64),
          "bcc4-512-block128-site-waves": (lambda: sy.heisenberg("bcc", 4), 512, None, 2, 4, 128)}[case]
    s, B, tile = mk[0](), mk[1], mk[2]
    R = mk[3] if len(mk) > 3 else None
    SP = mk[4] if len(mk) > 4 else 1
    blk = mk[5] if len(mk) > 5 else None
    g = G._Gen(s, G.MODE_F32, B, "uniform", replica_major=True, rm_tile=tile, rm_chunks=R, rm_sites=SP,
               rm_block=blk)
    mod = G.generate(s, G.MODE_F32, B, replica_major=True, rm_tile=tile, rm_chunks=R, rm_sites=SP, rm_block=blk)
    Gn, S, chunk = g.rm_plan()
    tx, ty, tz = g.rm_tile
    ntx, nty = -(-g.BX // tx), -(-g.L2 // ty)
    gx, gy, gz = mod.sweep[0].grid
    assert gx == G.XCDS and gy * gz >= chunk
    # replicas: each launch covers its block [rep_base, rep_base + block) through its groups (lanes past the
    # block's end skip the update), and the blocks cover every replica of the batch exactly once
    Rc = g.rm_chunks()
    per = 64 * Rc * (4 // SP)
    Bb = g.rm_block
    bases = sorted({l.rep_base for l in mod.sweep})
    assert bases == list(range(0, B, Bb)) and len(mod.sweep) == len(bases) * len(mod.classes)
    reps = np.concatenate([b0 + r for b0 in bases for r in [np.concatenate([grp * per + np.arange(per)
                                                                             for grp in range(Gn)])]
                           for r in [r[r < Bb]]])
    assert np.array_equal(np.sort(reps), np.arange(B))
    seen = np.zeros((Gn, g.L3, g.L2, g.BX), np.int64)
    for bx in range(gx):
        k = np.arange(gy * gz)
        v = bx * chunk + k[k < chunk]
        v = v[v < Gn * S]
        grp = np.repeat(v // S, SP)
        t = ((v % S)[:, None] * SP + np.arange(SP)[None, :]).reshape(-1) if SP > 1 else v % S
        tt, u = t // (tx * ty * tz), t % (tx * ty * tz)
        j = (tt % ntx) * tx + u % tx
        y = ((tt // ntx) % nty) * ty + (u // tx) % ty
        z = ((tt // ntx) // nty) * tz + u // (tx * ty)
        ok = (j < g.BX) & (y < g.L2) & (z < g.L3)
        np.add.at(seen, (grp[ok], z[ok], y[ok], j[ok]), 1)
    assert seen.min() == 1 and seen.max() == 1


def test_ctypes_structs_match_the_c_header(tmp_path):
    """Every field of the ctypes mirrors sits at the offset the C compiler gives it in include/sck.h, and the
    structs have the same size (a mismatch would pass the version check and corrupt every launch)."""
    import subprocess
    from sck_amd import _lib as L
    structs = {"sck_args": L.Args, "sck_launch": L.LaunchC, "sck_desc": L.Desc, "sck_mlp_sample": L.MlpSample,
               "sck_mlp_config": L.MlpConfig, "sck_mlp_adam_config": L.AdamConfig, "sck_width_grid": L.WidthGrid}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sck.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'    printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("    return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    inc = Path(__file__).resolve().parents[1] / "include"
    subprocess.run(["gcc", "-std=c11", "-I", str(inc), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_storage_modes_match_the_c_header():
    """The generator's storage modes carry the values of enum sck_mode in include/sck.h."""
    import re
    from sck_amd import hipgen as G
    hdr = (Path(__file__).resolve().parents[1] / "include" / "sck.h").read_text()
    enum = {m.group(1): int(m.group(2)) for m in re.finditer(r"SCK_MODE_(\w+) = (\d+)", hdr)}
    assert enum == {"INT8": G.MODE_INT8, "F32": G.MODE_F32, "PACKED": G.MODE_PACKED, "BITWISE": G.MODE_BITWISE,
                    "BITWISE_INDEP": G.MODE_BITWISE_INDEP, "F64": G.MODE_F64}


def test_float64_storage_runs_the_reference_layout():
    """float64 storage (updated in float32 arithmetic) generates the plain kernels with double loads and stores,
    stores the float32 spin back on rejection, and takes neither the replica-major, resident nor tiled kernels."""
    from sck_amd import hipgen as G
    from sck_amd.sampler import choose_layout
    s = sy.xy("triangular", 2048)
    assert choose_layout(s, G.MODE_F64, 2048) == ("reference", False)
    m = G.generate(sy.heisenberg("bcc", 4), G.MODE_F64, 4)
    assert m.resident is None and m.tiled is None and "sck_sample" in m.kernels
    upd = m.src[m.src.index("DEV int upd_0"):]
    upd = upd[:upd.index("\n}\n")]
    assert "const double *__restrict__ src" in upd and "} else {" in upd


@pytest.mark.parametrize("make, msg", [
    (lambda: sy.heisenberg("sc", 4), "Ising"),
    (lambda: sy.System("J1-J2", 1, cr.make("sc", (1, 2)), (4, 4, 4), (True,) * 3, {1: 1.0, 2: 0.5}), "uniform"),
    (lambda: sy.System("DMI", 1, cr.make("square"), (4, 4), (True,) * 2, {1: 1.0}, extra={0: np.eye(1)}), "extra"),
])
def test_bitwise_refuses_other_systems_with_the_reference_messages(make, msg):
    """The reference's refusals (its tests/test_bitwise.py cases): raised before any device is touched."""
    from sck_amd.sampler import Sampler
    with pytest.raises(ValueError, match=msg):
        Sampler(make(), 32, device="cpu", bitwise=True)


@pytest.mark.parametrize("kw, msg", [(dict(words_per_program=3, bitwise=True), "power of two"),
                                     (dict(max_block=48), "power of two"),
                                     (dict(independent=True), "bitwise=True"),
                                     (dict(packed=True), None)])
def test_sampler_argument_refusals(kw, msg):
    from sck_amd.sampler import Sampler
    s = sy.ising("square", 4)
    if msg is None:                              # packed XY is refused, packed Ising is fine (until the device)
        with pytest.raises(ValueError, match="Ising"):
            Sampler(sy.xy("square", 4), 32, device="cpu", **kw)
        return
    with pytest.raises(ValueError, match=msg):
        Sampler(s, 32, device="cpu", **kw)


def test_specialised_philox_equals_general_and_oracle(tmp_path):
    """philox10_z (counter (c0, c1, 0, 0), key (k0, 0); rounds 1-2 simplified) equals the general philox10 of
    the same generated header and the oracle's Philox-4x32-10, compiled for the host (bitop3 as its truth
    table), for c1 = 0 (32-bit counters) and for random high words (64-bit counters)."""
    import subprocess
    import oracle as O
    src = hg.HEADER
    body = src[src.index("DEV void philox10("):src.index("// uniform in [0, 1)")]
    rng = np.random.default_rng(5)
    c0 = np.concatenate([[0, 1, 0xFFFFFFFF], rng.integers(0, 2 ** 32, 200)]).astype(np.uint32)
    k0 = np.concatenate([[0, 0xFFFFFFFF, 7], rng.integers(0, 2 ** 31, 200)]).astype(np.uint32)
    c0, k0 = np.concatenate([c0, c0]), np.concatenate([k0, k0])
    c1 = np.concatenate([np.zeros(203, np.uint32), [1, 0xFFFFFFFF, 3], rng.integers(0, 2 ** 32, 200)]).astype(np.uint32)
    prog = ["#include <stdint.h>", "#include <stdio.h>", "#define DEV static inline",
            "static uint32_t __builtin_amdgcn_bitop3_b32(uint32_t a, uint32_t b, uint32_t c, unsigned t) {",
            "    uint32_t r = 0;",
            "    for (int i = 0; i < 32; ++i) {",
            "        unsigned idx = (((a >> i) & 1u) << 2) | (((b >> i) & 1u) << 1) | ((c >> i) & 1u);",
            "        r |= ((t >> idx) & 1u) << i;",
            "    }",
            "    return r;",
            "}", body, "int main(void) {",
            f"    static const uint32_t C0[{len(c0)}] = {{{', '.join(f'{v}u' for v in c0)}}};",
            f"    static const uint32_t K0[{len(k0)}] = {{{', '.join(f'{v}u' for v in k0)}}};",
            f"    static const uint32_t C1[{len(c1)}] = {{{', '.join(f'{v}u' for v in c1)}}};",
            f"    for (int i = 0; i < {len(c0)}; ++i) {{",
            "        uint32_t a0 = C0[i], a1 = C1[i], a2 = 0, a3 = 0, b0 = C0[i], b1 = C1[i], b2, b3;",
            "        philox10(a0, a1, a2, a3, K0[i], 0u);",
            "        philox10_z(b0, b1, b2, b3, K0[i]);",
            '        printf("%u %u %u %u %u %u %u %u\\n", a0, a1, a2, a3, b0, b1, b2, b3);',
            "    }", "    return 0;", "}"]
    cpp = tmp_path / "philox.cpp"
    cpp.write_text("\n".join(prog))
    exe = tmp_path / "philox"
    subprocess.run(["g++", "-std=c++17", "-O1", str(cpp), "-o", str(exe)], check=True)
    out = np.array([[int(v) for v in l.split()] for l in
                    subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
                    if l.strip()], dtype=np.uint64)
    assert np.array_equal(out[:, :4], out[:, 4:])
    for i in range(len(c0)):
        ref = O.philox((int(k0[i]), 0), [[int(c0[i]), int(c1[i]), 0, 0]])[0]
        assert np.array_equal(out[i, :4], ref.astype(np.uint64)), i


def test_replica_blocks_plan():
    """A replica-major lattice beyond the cache budget runs in replica blocks (C3: 32 blocks of 256 replicas,
    201 MB each) with the sweep list block-major; smaller lattices and blocks below one wave stay unblocked;
    a forced block must be a multiple of 64 dividing the batch."""
    s = sy.System("Heisenberg bcc J1J2", 3, cr.make("bcc", (1, 2)), (32, 32, 32), (True,) * 3, {1: 1.0, 2: -0.2},
                  h=(0.0, 0.0, 0.05), uniaxial=(0.1, (0.0, 0.0, 1.0)))
    m = hg.generate(s, hg.MODE_F32, 8192, replica_major=True)
    ncls = len(m.classes)
    assert m.rep_blocks == 32 and len(m.sweep) == 32 * ncls
    assert [l.rep_base for l in m.sweep] == [256 * (i // ncls) for i in range(len(m.sweep))]
    assert [l.kernel for l in m.sweep[:ncls]] == [l.kernel for l in m.sweep[ncls:2 * ncls]]
    assert 256 * s.nsite * 3 * 4 <= hg.RM_BLOCK_BYTES < 512 * s.nsite * 3 * 4
    assert hg.generate(s, hg.MODE_F32, 128, replica_major=True).rep_blocks == 1
    assert hg.generate(s, hg.MODE_F32, 8192, replica_major=True, rm_block=0).rep_blocks == 1
    assert hg.generate(sy.xy("triangular", 2048), hg.MODE_F32, 2048, replica_major=True).rep_blocks == 1
    small = sy.heisenberg("bcc", 4)
    f = hg.generate(small, hg.MODE_F32, 256, replica_major=True, rm_block=64)
    assert f.rep_blocks == 4 and "A.rep_base +" in f.src
    for bad in (32, 96, 100):
        with pytest.raises(ValueError, match="replica block"):
            hg.generate(small, hg.MODE_F32, 256, replica_major=True, rm_block=bad)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 prints a JSON error line and exits with status 2 (before any
    device use)."""
    import json
    import subprocess
    import sys
    r = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_reference_table_rows_cover_the_published_systems():
    """The systems bench.py's reference-table workload builds carry the names of the published rows, each with
    a rate per batch size (SCK/exp_perf.py:60-74; e6-throughput-h100.csv:2-14)."""
    import sys
    sys.path.insert(0, str(REPO))
    import bench
    names = [sy.ising("square", 8).name, sy.ising("triangular", 6).name, sy.ising("sc", 4).name,
             sy.ising("bcc", 4).name, sy.ising("fcc", 4).name, sy.xy("square", 8).name, sy.xy("sc", 4).name,
             sy.heisenberg("sc", 4).name, sy.heisenberg("bcc", 4).name, sy.heisenberg("fcc", 4).name,
             sy.heisenberg("triangular", 6, J=-1.0, label="Heisenberg triangular AFM").name,
             sy.heisenberg("kagome", 6, J=-1.0, label="Heisenberg kagome AFM").name, sy.nanodot(Rd=3.0, L=3).name]
    assert sorted(names) == sorted(bench.REF_H100)
    assert all(len(v) == len(bench.REF_H100_BATCHES) for v in bench.REF_H100.values())
    assert sorted(bench.REF_H100_SCAN) == [512, 1024, 2048, 4096]


def _field_by_sites(s, S):
    """The bond part of the local field, site by site: for every cell and bond, the neighbour's coordinates,
    wrapped on periodic axes and dropped beyond open faces, times the bond matrix."""
    x = np.asarray(S, np.float64)
    v = np.zeros_like(x)
    L = (s.L1, s.L2, s.L3)
    for i, bd in enumerate(s.crystal.bonds):
        M = s.bond_matrix(i)
        for z in range(s.L3):
            for y in range(s.L2):
                for xx in range(s.L1):
                    q = [xx, y, z]
                    inside = True
                    for k in range(s.crystal.D):
                        q[k] += bd.delta[k]
                        if s.periodic[k]:
                            q[k] %= L[k]
                        elif not 0 <= q[k] < L[k]:
                            inside = False
                    if inside:
                        v[:, :, bd.alpha, z, y, xx] += x[:, :, bd.alpha2, q[2], q[1], q[0]] @ M.T
    return v


@pytest.mark.parametrize("make", [
    lambda: sy.heisenberg("bcc", 3),
    lambda: sy.System("Heis fcc open", 3, cr.make("fcc"), (3, 2, 4), (False, True, False), {1: 1.0, 2: -0.3}),
    lambda: sy.xy("triangular", 5),
    lambda: sy.ising("square", 5),
], ids=["bcc", "fcc_open", "triangular", "square"])
def test_local_field_equals_the_site_by_site_sum(make):
    """Sampler.local_field (the float64 check quantity of SCK/sampler.py:296-324, here an index gather) against
    the same sum written site by site; it runs on the host, so it takes a stand-in holding only the system."""
    import types

    import torch

    from sck_amd.sampler import Sampler
    s = make()
    rng = np.random.default_rng(4)
    S = rng.standard_normal((2, s.n) + tuple(s.shape))
    fake = types.SimpleNamespace(sys=s, unpack=lambda X: X)
    fake._bond_gather = types.MethodType(Sampler._bond_gather, fake)
    v = Sampler.local_field(fake, torch.as_tensor(S)).numpy()
    np.testing.assert_allclose(v, _field_by_sites(s, S), rtol=0, atol=1e-12)
