"""Packer checks on the CPU: packed device code, run through a CPU model of
the interpreter, must reproduce the oracle's residues and FP64 matrices."""

import numpy as np
import pytest

from conftest import program_from_json
import devmodel

from dreams_rns import _native as N
from dreams_rns.compiler import CmfProgram, Instruction, Opcode, compile_cmf_from_dict
from oracle import cref
from oracle import walk as ow

PRIMES = [int(p) for p in ow.generate_rns_primes(4)]


def _oracle_mf(program, axes, depth):
    out = []
    for s in range(depth):
        axis_f = [float(sn * td + s * tn * sd) / float(sd * td) for sn, sd, tn, td in axes]
        out.append(ow.eval_float(ow._ops(program), program.m, s, axis_f, program.constants).reshape(-1))
    return out


def _check(program, axes, depth, init_res=None, init_f=None):
    m = program.m
    words, slots, meta = N.pack_program(program)
    if init_res is None:
        init_res = np.zeros((m * m, len(PRIMES)), dtype=np.int64)
        init_f = np.zeros(m * m)
        for d in range(m):
            init_res[d * m + d] = 1
            init_f[d * m + d] = 1.0
    res, mfs = devmodel.walk(program, words, meta, slots, PRIMES, axes, depth, init_res, init_f)
    ref_res, _, _ = cref.walk(program, depth, PRIMES, axes, init_res, init_f)
    assert np.array_equal(np.array(res, dtype=np.int64), ref_res)
    for a, b in zip(mfs, _oracle_mf(program, axes, depth)):
        assert [float(x).hex() for x in a] == [float(x).hex() for x in b]
    return meta


def test_golden_pcf_programs(golden):
    for c in golden["compile_pcf"]:
        if c["program"] is None:
            continue
        p = program_from_json(c["program"])
        _check(p, [(1, 1, 1, 1)], 6)


def test_golden_cmf_programs(golden):
    for c in golden["run_cmf_walk"]:
        p = program_from_json(c["program"])
        sh = c["shift"]
        tr = c["traj"] or [[d, 1] for d in c["directions"]]
        axes = [(sh[i][0], sh[i][1], tr[i][0], tr[i][1]) for i in range(c["dim"])]
        _check(p, axes, 4)


def _prog(m, dim, ops, consts, dirs=None):
    return CmfProgram(m, dim, [Instruction(Opcode(o[0]), *o[1:]) for o in ops] + [Instruction(Opcode.END)],
                      consts, dirs or [1] * dim)


def test_dead_code_unwritten_registers_and_overwrites():
    # r5 is never written (reads 0); entry (0,0) stored twice (last wins);
    # an unused INV is dead; (1,0) stores a constant 7; (0,1) stores the axis.
    ops = [(1, 0, 0), (2, 0, 1), (6, 0, 1, 2), (10, 2, 3), (4, 2, 5, 4), (11, 4, 0, 0), (11, 1, 0, 0),
           (11, 1, 1, 0), (11, 0, 0, 1), (8, 0, 6), (11, 6, 1, 1)]
    p = _prog(2, 1, ops, [7])
    meta = _check(p, [(3, 1, 2, 1)], 5)
    assert meta["kinds"] & 3 == 2


def test_zero_and_one_entries_are_static():
    p = compile_cmf_from_dict({(0, 0): "0", (0, 1): "x", (1, 0): "1", (1, 1): "x**2 + 1"}, 2, 1, ["x"])
    meta = _check(p, [(1, 1, 1, 1)], 5)
    assert meta["kinds"] & 3 == 0 and (meta["kinds"] >> 4) & 3 == 1


def test_long_sums_insert_reduction():
    # 2^12 chained additions of a large value exceed the 2^40 bound: RED is inserted
    ops = [(1, 0, 0), (4, 0, 0, 1)]
    for i in range(2048):
        ops.append((4, 1 + i % 2, 0, 2 - i % 2))   # r1 <-> r2 ping-pong: acc += x
    ops.append((6, 1, 1, 3))
    ops.append((11, 3, 1, 1))
    p = _prog(2, 1, ops, [])
    words, _, _ = N.pack_program(p)
    assert any((int(w) & 255) == devmodel.D_RED for w in words[::2])
    _check(p, [(123456789, 1, 987654321, 1)], 2)


def test_many_live_values_spill_to_lds_slots():
    # 24 values live at once force LDS temp slots
    ops = [(1, 0, 0)]
    for i in range(1, 25):
        ops.append((8, 0, i))          # r_i = x^2 (24 live values)
    acc = 1
    for i in range(2, 25):
        ops.append((4, acc, i, 100 + i))
        acc = 100 + i
    ops.append((11, acc, 1, 1))
    p = _prog(2, 1, ops, [])
    _, _, meta = N.pack_program(p)
    assert meta["nslots"] > 0
    _check(p, [(2, 1, 3, 1)], 3)


def test_many_constants_go_to_lds():
    consts = list(range(3, 3 + 40))
    ops = [(2, 0, 0)]
    acc = 0
    for i in range(1, 40):
        ops.append((2, i, 200 + i))
        ops.append((6, acc, 200 + i, 300 + i))
        acc = 300 + i
    ops.append((11, acc, 0, 1))
    p = _prog(2, 1, ops, consts)
    _, _, meta = N.pack_program(p)
    assert meta["const_lds"] == 1
    _check(p, [(1, 1, 1, 1)], 2)


def test_packer_rejects_bad_programs():
    with pytest.raises(ValueError):
        N.pack_program(_prog(2, 1, [(1, 3, 0)], []))        # axis out of range
    with pytest.raises(ValueError):
        N.pack_program(_prog(2, 1, [(11, 0, 2, 0)], []))    # row out of range
    with pytest.raises(ValueError):
        N.pack_program(_prog(5, 1, [], []))                 # m > 4
    with pytest.raises(ValueError):
        N.pack_program(_prog(2, 1, [(2, 5, 0)], [1]))       # constant index
