"""Host-side checks (CPU only): compiler, FP64 residue arithmetic formulas,
the C ABI's exported symbols, workload generators, API surface."""

import math
import random
import re
from fractions import Fraction

import numpy as np
import pytest

from conftest import ROOT, program_from_json

import dreams_rns as D
from dreams_rns import _native as N
from dreams_rns import synth
from dreams_rns.compiler import CmfProgram, compile_cmf_from_dict
from oracle import exact
from oracle import walk as ow


# ------------------------------------------------------------------ compiler
def test_compile_pcf_bytecode_identity(golden):
    for c in golden["compile_pcf"]:
        p = D.compile_pcf(c["a"], c["b"])
        if c["program"] is None:
            assert p is None
            continue
        assert [[int(i.op), i.arg0, i.arg1, i.arg2] for i in p.opcodes] == c["program"]["ops"], c["a"]
        assert [str(v) for v in p.constants] == c["program"]["constants"]
        assert D.pcf_initial_values(c["a"]) == c["a0"]


def test_compile_cmf_bytecode_identity(golden):
    for c in golden["run_cmf_walk"]:
        mat = {(a, b): e for a, b, e in c["matrix"]}
        p = compile_cmf_from_dict(mat, c["m"], c["dim"], c["axes"], c["directions"])
        assert [[int(i.op), i.arg0, i.arg1, i.arg2] for i in p.opcodes] == c["program"]["ops"]


def test_pcf_initial_values_readme():
    assert D.pcf_initial_values("2") == 2  # README.md:44


def test_compiled_program_equals_sympy_value():
    import sympy as sp

    rng = random.Random(5)
    x, y = sp.symbols("x y")
    for _ in range(40):
        terms = []
        for _ in range(rng.randint(1, 4)):
            c = sp.Rational(rng.randint(-9, 9) or 1, rng.choice([1, 1, 2, 3]))
            terms.append(c * x ** rng.randint(0, 5) * y ** rng.randint(0, 3))
        expr = sp.Add(*terms)
        if rng.random() < 0.5:
            expr = expr / (x + rng.randint(1, 4))
        p = compile_cmf_from_dict({(0, 1): str(expr), (1, 0): "1"}, 2, 2, ["x", "y"])
        for xv, yv in ((2, 3), (Fraction(1, 3), -2), (5, Fraction(7, 2))):
            M = exact.eval_exact(p, 0, [Fraction(xv), Fraction(yv)])
            assert M[0][1] == Fraction(sp.Rational(expr.subs({x: sp.Rational(xv), y: sp.Rational(yv)})))
            assert M[1][0] == 1 and M[0][0] == 0


def test_compiler_errors_like_reference():
    import sympy as sp

    with pytest.raises(ValueError):
        compile_cmf_from_dict({(0, 0): "z"}, 2, 1, ["x"])          # unknown symbol
    with pytest.raises(ValueError):
        compile_cmf_from_dict({(0, 0): "sqrt(x)"}, 2, 1, ["x"])    # unsupported power
    with pytest.raises(ValueError):
        compile_cmf_from_dict({(0, 0): "sin(x)"}, 2, 1, ["x"])     # unsupported type
    comp = D.CmfCompiler(2, 1)
    with pytest.raises(RuntimeError):
        for _ in range(600):
            comp.load_n()


def test_serialize_round_trip(golden):
    for c in golden["run_cmf_walk"]:
        p = program_from_json(c["program"])
        q = CmfProgram.deserialize(p.serialize())
        assert (q.m, q.dim, q.constants, q.directions) == (p.m, p.dim, p.constants, p.directions)
        assert [(int(i.op), i.arg0, i.arg1, i.arg2) for i in q.opcodes] == [
            (int(i.op), i.arg0, i.arg1, i.arg2) for i in p.opcodes]


def test_serialize_byte_layout():
    # the layout compiler.py:53-62 writes, spelled out field by field in
    # explicit little-endian: header (m, dim, n_ops, n_const) as u32, then
    # 16 bytes (op, a0, a1, a2) per instruction, int64 constants, int32
    # directions
    import struct

    from dreams_rns.compiler import Instruction, Opcode

    p = CmfProgram(m=2, dim=2, opcodes=[Instruction(Opcode.LOAD_X, 0, 0, 0), Instruction(Opcode.LOAD_C, 1, 1, 0),
                                        Instruction(Opcode.MUL, 2, 0, 1), Instruction(Opcode.STORE, 2, 3, 0),
                                        Instruction(Opcode.END, 0, 0, 0)],
                   constants=[-7, 2 ** 62], directions=[3, -1])
    want = bytes.fromhex("02000000" "02000000" "05000000" "02000000")
    # opcode numbers of compiler.py:14-28: LOAD_X 1, LOAD_C 2, MUL 6, STORE 11, END 12
    for op, a0, a1, a2 in [(1, 0, 0, 0), (2, 1, 1, 0), (6, 2, 0, 1), (11, 2, 3, 0), (12, 0, 0, 0)]:
        want += b"".join(x.to_bytes(4, "little") for x in (op, a0, a1, a2))
    want += (-7).to_bytes(8, "little", signed=True) + (2 ** 62).to_bytes(8, "little", signed=True)
    want += (3).to_bytes(4, "little", signed=True) + (-1).to_bytes(4, "little", signed=True)
    got = p.serialize()
    assert got == want
    assert len(got) == 16 + 16 * 5 + 8 * 2 + 4 * 2
    q = CmfProgram.deserialize(got)
    assert (q.constants, q.directions) == ([-7, 2 ** 62], [3, -1])
    # a constant past int64 cannot be written ('q' format), as in the reference
    with pytest.raises(struct.error):
        CmfProgram(m=2, dim=0, opcodes=[], constants=[2 ** 63]).serialize()


# ----------------------------------------------- FP64 residue arithmetic
def _mm(a: float, b: float, p: float, pinv: float) -> float:
    """The device's balanced modular multiply, step by step in IEEE doubles."""
    h = a * b
    l = exact.fma(a, b, -h)
    q = float(round(h * pinv))  # round-half-even, like v_rndne_f64
    r = exact.fma(-q, p, h)
    return r + l


def test_fp64_modmul_is_exact_within_bounds():
    rng = random.Random(11)
    primes = [int(v) for v in ow.generate_rns_primes(128)]
    for _ in range(3000):
        p = rng.choice(primes)
        pd, pinv = float(p), 1.0 / float(p)
        lim = rng.choice([p // 2, p, 1 << 33, 1 << 40])
        a, b = rng.randint(-lim, lim), rng.randint(-lim, lim)
        r = _mm(float(a), float(b), pd, pinv)
        assert r == int(r) and (int(r) - a * b) % p == 0
        assert abs(r) <= p / 2 + 1.5 * 2.0 ** -52 * abs(a * b) + 1
    # multiples of p reduce to exactly zero (INV(0) = 0 through Fermat)
    for p in primes[:8]:
        assert _mm(float(p), float(p), float(p), 1.0 / p) == 0.0
        assert _mm(0.0, 12345.0, float(p), 1.0 / p) == 0.0


def test_fp64_fermat_inverse_formula():
    rng = random.Random(2)
    for p in [int(v) for v in ow.generate_rns_primes(8)]:
        pd, pinv = float(p), 1.0 / p
        for a in [0, 1, p - 1, rng.randint(2, p - 2), -(p // 3)]:
            r = float(a)
            e = p - 2
            for bit in range(29, -1, -1):
                r = _mm(r, r, pd, pinv)
                t = _mm(r, float(a), pd, pinv)
                r = t if (e >> bit) & 1 else r
            assert int(r) % p == pow(a % p, p - 2, p)


# -------------------------------------------------------------------- C ABI
def test_library_exports_every_declared_symbol():
    header = open(f"{ROOT}/include/dreams_hip.h").read()
    declared = set(re.findall(r"\b(drx_[a-z_]+)\s*\(", header))
    lib = N.load_library()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert set(N.SIGNATURES) >= declared
    assert lib.drx_abi_version() == N.ABI_VERSION


def test_struct_layouts_match_header():
    import ctypes as C

    assert C.sizeof(N.Walk) == 72 and N.WALK_DTYPE.itemsize == 72
    assert N.Walk.mode.offset == 20 and N.Walk.out_p.offset == 64 and N.Walk.out_q.offset == 68
    assert N.WALK_DTYPE.fields["mode"][1] == 20 and N.WALK_DTYPE.fields["out_q"][1] == 68
    assert C.sizeof(N.WalkOut) == 48 and N.OUT_DTYPE.itemsize == 48
    assert C.sizeof(N.ServerResult) == 8 + 48 + 8 + 24 + 8


def test_struct_layouts_match_the_c_compiler(tmp_path):
    # sizes and offsets of every ctypes mirror against gcc on the header itself
    import ctypes as C
    import subprocess

    mirrors = {"drx_walk": N.Walk, "drx_walk_out": N.WalkOut, "drx_batch": N.Batch, "drx_launch_cfg": N.LaunchCfg,
               "drx_server_cfg": N.ServerCfg, "drx_server_result": N.ServerResult}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dreams_hip.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'  printf("{cname}.{f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT}/include", str(src), "-o", str(exe)], check=True)
    got = dict(ln.split() for ln in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for cname, py in mirrors.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f in py._fields_:
            assert int(got[f"{cname}.{f[0]}"]) == getattr(py, f[0]).offset, (cname, f[0])


# -------------------------------------------------------------- generators
def test_generators_are_seeded_and_shaped():
    a = synth.pcf2(n_walks=64)
    b = synth.pcf2(n_walks=64)
    assert a == b and len(a["walks"]) == 64 and a["depth"] == 2000 and a["K"] == 32
    for w in a["walks"]:
        assert -5.0 <= w["L"] <= 5.0
        D.compile_pcf(w["a"], w["b"])
    c1 = synth.pcf1()
    assert c1["depth"] == 2000 and c1["K"] == 32
    rng = np.random.default_rng(1)
    cf = lambda mat, m, dim, axes, dirs: compile_cmf_from_dict(mat, m, dim, axes, dirs)
    spec = synth.cmf_walk_spec(rng, 3, cf, 100, 250, True)
    assert 100 <= len(spec["program"].opcodes) <= 250
    assert any(int(i.op) == 10 for i in spec["program"].opcodes)
    assert any(spec["directions"]) and all(0 <= s <= 5 for s in spec["shift"])
    assert 16 <= synth.k_for(synth.estimate_bits(6, 20, 500)) <= 128


def test_public_api_surface():
    for name in ("compile_cmf", "compile_cmf_from_dict", "CmfCompiler", "CmfProgram", "Opcode", "WalkConfig",
                 "generate_rns_primes", "crt_reconstruct", "centered", "compute_dreams_delta_float",
                 "compute_dreams_delta_exact", "compile_pcf", "pcf_initial_values", "run_pcf_walk", "verify_pcf",
                 "run_cmf_walk", "gpu_available", "run_pcf_walk_gpu", "run_pcf_walk_batch_gpu"):
        assert hasattr(D, name), name


def test_no_cpu_fallback_without_device():
    # this container has no HIP device: the product path must raise, not compute
    if D.gpu_available():
        pytest.skip("a HIP device is present")
    p = D.compile_pcf("2", "n**2")
    with pytest.raises(N.NativeUnavailable):
        D.run_pcf_walk(p, 2, 10, 8)
