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


def _res_code(words, meta):
    return devmodel.decode_res(words[: meta["words"]])


def _fp_code(words, meta):
    fw = meta["words"] + (meta["words"] & 1)
    return devmodel.decode_res(words[fw:fw + meta["fwords"]])


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
    words, _, meta = N.pack_program(p)
    assert any(c[0] == devmodel.RO_RED for c in _res_code(words, meta))
    _check(p, [(123456789, 1, 987654321, 1)], 2)


def test_many_live_values_spill_to_lds_slots():
    # 24 values, each used by a forward sum (entry (0,1)) and by a backward
    # product (entry (1,1)): all are live at once in any evaluation order, so
    # both codes need LDS temp slots
    ops = [(1, 0, 0)]
    for i in range(1, 25):
        ops.append((4, 0, 0, i) if i == 1 else (4, i - 1, 0, i))   # r_i = i * x, distinct values
    acc = 1
    for i in range(2, 25):
        ops.append((4, acc, i, 100 + i))
        acc = 100 + i
    ops.append((11, acc, 0, 1))
    acc = 24
    for i in range(23, 0, -1):
        ops.append((6, acc, i, 200 + i))
        acc = 200 + i
    ops.append((11, acc, 1, 1))
    p = _prog(2, 1, ops, [])
    _, _, meta = N.pack_program(p)
    assert meta["nslots"] > 0 and meta["fnslots"] > 0
    _check(p, [(2, 1, 3, 1)], 3)


def test_many_constants_go_to_lds():
    # x * c0 * c1 * ... * c39: 40 constants referenced by step-dependent work
    consts = list(range(3, 3 + 40))
    ops = [(1, 0, 500), (2, 0, 0), (6, 500, 0, 501)]
    acc = 501
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


def test_constant_subexpressions_fold():
    # c0 * c1, INV(c2), -(c0 + c1)^2 and the INV of the zero constant c3 are
    # step-independent: folded into constants computed at upload; INV(c2)
    # alone no longer blocks forward differences
    ops = [(2, 0, 0), (2, 1, 1), (2, 2, 2), (2, 3, 3), (1, 0, 4),
           (6, 0, 1, 10), (10, 2, 11), (4, 0, 1, 12), (8, 12, 13), (7, 13, 14), (10, 3, 15),
           (6, 4, 10, 20), (4, 20, 11, 21), (11, 21, 0, 1),          # (0,1) = x c0 c1 + 1/c2
           (4, 14, 15, 22), (6, 22, 4, 23), (11, 23, 1, 1),          # (1,1) = x (-(c0+c1)^2 + INV(0))
           (11, 11, 1, 0)]                                            # (1,0) = 1/c2
    p = _prog(2, 1, ops, [5, -7, 3, 0])
    meta = _check(p, [(2, 3, 5, 7)], 4)
    ops_folded = {d[0] for d in meta["derived"]}
    assert {devmodel.D_MUL, devmodel.D_INV, devmodel.D_ADD, devmodel.D_SQR, devmodel.D_NEG} <= ops_folded
    assert meta["fd_ok"] == 1 and meta["fd_degree"] == 1
    words, _, meta = N.pack_program(p)
    assert all(c[0] != devmodel.RO_INV for c in _res_code(words, meta))


def _inv_program(nested):
    # 2x2 program with four step-dependent INVs: 1/(x - 2), 1/(x - 5),
    # 1/(x^2 + 1), 1/(3x); with `nested`, the last one is 1/(1/(x-2) + x)
    ops = [(1, 0, 0), (2, 0, 1), (2, 1, 2), (2, 2, 3), (2, 3, 4),
           (5, 0, 1, 10), (10, 10, 11),                    # 1/(x-2)
           (5, 0, 2, 12), (10, 12, 13),                    # 1/(x-5)
           (8, 0, 14), (4, 14, 3, 15), (10, 15, 16)]       # 1/(x^2+1)
    if nested:
        ops += [(4, 11, 0, 17), (10, 17, 18)]              # 1/(1/(x-2) + x)
    else:
        ops += [(6, 0, 4, 17), (10, 17, 18)]               # 1/(3x)
    ops += [(6, 11, 13, 20), (11, 20, 0, 0), (11, 16, 0, 1), (4, 18, 13, 21), (11, 21, 1, 0),
            (5, 11, 16, 22), (11, 22, 1, 1)]
    return _prog(2, 1, ops, [2, 5, 1, 3])


def test_independent_inversions_share_one():
    # steps 0..6 pass x = 2 and x = 5 (zero arguments: INV(0) = 0) and x = 0
    # (1/(3x) of zero) through the shared inversion
    p = _inv_program(False)
    meta = _check(p, [(0, 1, 1, 1)], 7)
    assert meta["inv_batched"] == 4
    words, _, meta = N.pack_program(p)
    assert sum(1 for c in _res_code(words, meta) if c[0] == devmodel.RO_INV) == 1
    assert sum(1 for c in _fp_code(words, meta) if c[0] == devmodel.FO_INV) == 4


def test_dependent_inversion_stays_separate():
    p = _inv_program(True)
    meta = _check(p, [(0, 1, 1, 1)], 7)
    assert meta["inv_batched"] == 3
    words, _, meta = N.pack_program(p)
    assert sum(1 for c in _res_code(words, meta) if c[0] == devmodel.RO_INV) == 2


def test_vector_packing_runs_the_state_walk():
    # vector-packed odd-zeta programs (4x4 / 6x6 / 8x8, 6 / 9 / 12 axes):
    # the code computes v' = M(s) v; through the device model it equals the
    # oracle's state-vector walk in residues and, within rounding, in FP64
    from fractions import Fraction
    import odd_zeta as oz

    fx = oz.load()
    for name, spec in fx["specs"].items():
        prog = oz.compile_spec(spec)
        words, slots, meta = N.pack_program(prog, N.DRX_WALK_VECTOR)
        v0 = oz.initial_state(spec["n"])
        shift = oz.parse_vec("[2,3,2,3,1,2,3,1,2,3,1,2]")[: spec["dim"]]
        traj = [Fraction(1, 3)] + [Fraction(d) for d in spec["directions"][1:]]
        axes = [(s.numerator, s.denominator, t.numerator, t.denominator) for s, t in zip(shift, traj)]
        v0_res = [[(x.numerator % p) * pow(x.denominator, p - 2, p) % p for p in PRIMES] for x in v0]
        res, vf, ls = devmodel.walk_vec(prog, words, meta, slots, PRIMES, axes, 7, v0_res, [float(x) for x in v0],
                                        spec["accumulator_idx"], spec["constant_idx"])
        ref = ow.run_cmf_walk_vec(prog, 7, len(PRIMES), shift, v0, spec["accumulator_idx"], spec["constant_idx"],
                                  traj)
        assert res[spec["accumulator_idx"]] == ref["p_residues"].tolist(), name
        assert res[spec["constant_idx"]] == ref["q_residues"].tolist(), name
        for got, want in ((vf[spec["accumulator_idx"]], ref["p_float"]), (vf[spec["constant_idx"]], ref["q_float"])):
            assert abs(got - want) <= 1e-13 * abs(want), (name, got, want)
        assert abs(ls - ref["log_scale"]) <= 1e-12 * max(1.0, abs(ls)), name


def test_vector_packing_limits():
    p = compile_cmf_from_dict({(0, 0): "x", (1, 1): "1"}, 2, 1, ["x"])
    N.pack_program(p, N.DRX_WALK_VECTOR)
    big = compile_cmf_from_dict({(i, i): "x + %d" % i for i in range(17)}, 17, 1, ["x"])
    with pytest.raises(ValueError):
        N.pack_program(big, N.DRX_WALK_VECTOR)        # 17 state entries
    with pytest.raises(ValueError):
        N.pack_program(p, 5)                           # unknown mode
