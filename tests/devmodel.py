"""CPU model of the packed device code (test helper).

Executes the words produced by the packer (drx_program_pack) with the bank
semantics of the device interpreters, using exact integers mod p
for the residue code and Python floats for the FP64 code, then the per-step
product P <- P * M exactly like the walk kernel's entry-kind matmul. Folded
constants are computed from the program's constants the way the upload does.
Agreement with the oracle validates the packer's constant folding, shared
inversion, dead-code removal, operand routing and register allocation on the
CPU.
"""

from __future__ import annotations

from typing import List, Sequence

D_ADD, D_SUB, D_MUL, D_NEG, D_SQR, D_CUB, D_INV, D_RED, D_MOV, D_NZ1, D_ZSEL = range(1, 12)
FO_NEG, FO_CUB, FO_INV, FO_MOV, FO_ZERO = range(3, 8)
STEP_SLOT = 15  # the model's axis list X: axis i at X[i], the step at X[15]


def _cube(x: float) -> float:
    from fractions import Fraction

    return float(Fraction(x) ** 3)


LAZY_LIMIT = 1 << 40


def _bal(v: int, p: int) -> int:
    return (v + p // 2) % p - p // 2


(RO_MUL, RO_ADD, RO_LIN, RO_MULNZ, RO_ZSEL, RO_INV, RO_RED, RO_NZ1, RO_LDM, RO_LDC, RO_LDV, RO_STM, RO_STE) = range(13)
RO_NOP = 15
RES_STEP = 15
RES_TEMPS, RES_CBASE = 16, 16


def decode_res(words):
    """Residue instructions: (op, x, y, dst, aux) with x, y bank indices
    (temps 0..15, input slot j at 16 + j), dst a bank index and aux the LDS
    row or the LIN signs."""
    return [(w & 15, (w >> 4) & 31, (w >> 9) & 31, (w >> 14) & 31, w >> 19) for w in (int(v) for v in words)]


def _sext2(v):
    v &= 3
    return v - 4 if v & 2 else v


def run_res(code, C, lc, p, nrows, state=None):
    """One step of the residue interpreter with lazy residues as on the
    device: ADD / LIN results stay unreduced, products and reductions are
    balanced; a multiply (or inversion) operand beyond 2^40 would break the
    device's exact FP64 product, so the packer must have inserted RED before
    that. C holds the input slots (step, axes, first constant slots), lc the
    further constant slots. Every instruction writes its dst; STM, STE and
    NOP must carry dst = x and never touch an input slot. Returns the entry
    bank E (16 slots)."""
    B = [0] * RES_TEMPS + [int(c) for c in C]
    E = [None] * 16
    M = [0] * max(1, nrows)
    for op, x, y, d, aux in code:
        xv, yv = B[x], B[y]
        if op == RO_MUL:
            assert abs(xv) <= LAZY_LIMIT and abs(yv) <= LAZY_LIMIT, "operand beyond 2^40"
            r = _bal(xv * yv, p)
        elif op == RO_ADD:
            r = xv + yv
        elif op == RO_LIN:
            r = _sext2(aux >> 2) * yv + _sext2(aux) * xv
        elif op == RO_MULNZ:
            assert abs(xv) <= LAZY_LIMIT and abs(yv) <= LAZY_LIMIT, "operand beyond 2^40"
            r = _bal(xv * (yv if yv % p else 1), p)
        elif op == RO_ZSEL:
            r = 0 if yv % p == 0 else xv
        elif op == RO_INV:
            assert abs(xv) <= LAZY_LIMIT, "operand beyond 2^40"
            r = _bal(pow(xv % p, p - 2, p), p)
        elif op == RO_RED:
            r = _bal(xv, p)
        elif op == RO_NZ1:
            r = xv if xv % p else 1
        elif op == RO_LDM:
            r = M[aux]
        elif op == RO_LDC:
            r = lc[aux]
        elif op == RO_LDV:
            r = state[aux]
        elif op in (RO_STM, RO_STE, RO_NOP):
            assert d == x, "a store or no-op must carry dst = x"
            if op == RO_STM:
                M[aux] = xv
            elif op == RO_STE:
                E[aux] = xv
            r = xv
        else:
            raise AssertionError(f"unknown residue op {op}")
        if op not in (RO_STM, RO_STE, RO_NOP):
            assert d < RES_TEMPS, "a result written to an input slot"
        B[d] = r
    return E


def run_fp(code, X, consts, dim, nslots: int = 0, V=None):
    """One step of the FP64 interpreter (same encoding as the residue code,
    FP64 operations): bank B holds 16 temps and the input slots (the first
    15 - dim constants, axis i at slot 14 - i, the step at slot 15); LDC reads
    constant slot 15 - dim + row. X: axis values then X[15] = step. Returns
    the entry bank E (16 slots)."""
    nc = RES_STEP - dim
    C = [0.0] * 16
    for i in range(min(nc, len(consts))):
        C[i] = consts[i]
    for ax in range(dim):
        C[RES_STEP - 1 - ax] = X[ax]
    C[RES_STEP] = X[STEP_SLOT]
    B = [0.0] * RES_TEMPS + C
    E = [0.0] * 16
    M = [0.0] * max(1, nslots)
    for op, x, y, d, aux in code:
        xv, yv = B[x], B[y]
        if op == RO_MUL:
            r = xv * yv
        elif op == RO_ADD:
            r = xv + yv
        elif op == RO_LIN:
            sx, sy = _sext2(aux), _sext2(aux >> 2)
            assert (sx, sy) == (1, -1), "FP64 LIN is x - y"
            r = xv - yv
        elif op == FO_NEG:
            r = -xv
        elif op == FO_CUB:
            r = _cube(xv)
        elif op == FO_INV:
            r = 1.0 / xv if abs(xv) > 1e-300 else 0.0
        elif op == FO_MOV:
            r = xv
        elif op == FO_ZERO:
            r = 0.0
        elif op == RO_LDM:
            r = M[aux]
        elif op == RO_LDC:
            r = consts[nc + aux]
        elif op == RO_LDV:
            r = V[aux]
        elif op in (RO_STM, RO_STE, RO_NOP):
            assert d == x, "a store or no-op must carry dst = x"
            if op == RO_STM:
                M[aux] = xv
            elif op == RO_STE:
                E[aux] = xv
            r = xv
        else:
            raise AssertionError(f"unknown FP64 op {op}")
        if op not in (RO_STM, RO_STE, RO_NOP):
            assert d < RES_TEMPS, "a result written to an input slot"
        B[d] = r
    return E


def all_constants(program, meta, p=None):
    """The program's constants followed by the folded ones: exact residues
    mod p, or FP64 values when p is None."""
    if p is None:
        vals = [float(int(c)) for c in program.constants]
    else:
        vals = [int(c) % p for c in program.constants]
    for op, a, b in meta["derived"]:
        x, y = vals[a], (vals[b] if b >= 0 else None)
        if p is None:
            v = {D_ADD: lambda: x + y, D_SUB: lambda: x - y, D_MUL: lambda: x * y, D_NEG: lambda: -x,
                 D_SQR: lambda: x * x, D_INV: lambda: (1.0 / x if abs(x) > 1e-300 else 0.0)}[op]()
        else:
            v = {D_ADD: lambda: (x + y) % p, D_SUB: lambda: (x - y) % p, D_MUL: lambda: x * y % p,
                 D_NEG: lambda: -x % p, D_SQR: lambda: x * x % p, D_INV: lambda: pow(x, p - 2, p)}[op]()
        vals.append(v)
    return vals


def walk(program, words, meta, const_slots, primes, axes, depth, init_res, init_f):
    """Residues (m, m, K) and M_f list per step of a walk through the model."""
    m, dim = program.m, program.dim
    code = decode_res(words[: meta["words"]])
    fw = meta["words"] + (meta["words"] & 1)  # the FP64 code starts on an even word
    fcode = decode_res(words[fw:fw + meta["fwords"]])
    kinds = meta["kinds"]
    K = len(primes)
    P = [[[int(init_res[i * m + j][k]) for j in range(m)] for i in range(m)] for k in range(K)]
    for k, p in enumerate(primes):
        p = int(p)
        cp = all_constants(program, meta, p)
        consts = [cp[ci] for ci in const_slots]
        n_cbank = RES_STEP - dim
        lc = consts[n_cbank:]
        for s in range(depth):
            C = [0] * 16
            for i in range(min(n_cbank, len(consts))):
                C[i] = consts[i]
            for ax in range(dim):
                sn, sd, tn, td = axes[ax]
                num, den = sn * td + s * tn * sd, sd * td
                C[RES_STEP - 1 - ax] = num % p * pow(den % p, p - 2, p) % p if den != 1 else num % p
            C[RES_STEP] = s
            E = run_res(code, C, lc, p, meta["nslots"])
            for e in range(m * m):
                if (kinds >> (2 * e)) & 3 == 2:
                    assert E[e] is not None, "general entry never stored"
                    assert abs(E[e]) <= LAZY_LIMIT, "entry beyond 2^40"
            C = [[0] * m for _ in range(m)]
            for i in range(m):
                for j in range(m):
                    acc = 0
                    for t in range(m):
                        kd = (kinds >> (2 * (t * m + j))) & 3
                        if kd == 2:
                            acc += P[k][i][t] * E[t * m + j]
                        elif kd == 1:
                            acc += P[k][i][t]
                    C[i][j] = acc % p
            P[k] = C
    res = [[[P[k][i][j] for k in range(K)] for j in range(m)] for i in range(m)]
    mfs = []
    cf = all_constants(program, meta)
    fconsts = [cf[ci] for ci in const_slots]
    for s in range(depth):
        X = [0.0] * 16
        for ax in range(dim):
            sn, sd, tn, td = axes[ax]
            X[ax] = float(sn * td + s * tn * sd) / float(sd * td)
        X[STEP_SLOT] = float(s)
        E = run_fp(fcode, X, fconsts, dim, meta["fnslots"])
        mf = []
        for e in range(m * m):
            kd = (kinds >> (2 * e)) & 3
            mf.append(E[e] if kd == 2 else (1.0 if kd == 1 else 0.0))
        mfs.append(mf)
    return res, mfs


def walk_vec(program, words, meta, const_slots, primes, axes, depth, v0_res, v0_f, acc_idx, const_idx):
    """A state-vector walk through the model of the vector-packed code: the
    residue code reads the state with LDV and stores the new state as its
    entries; the FP64 code reads bank V; the shadow rescales by max |v_f|
    above 1e10 (NaN: no rescale)."""
    import math

    m, dim = program.m, program.dim
    code = decode_res(words[: meta["words"]])
    fw = meta["words"] + (meta["words"] & 1)
    fcode = decode_res(words[fw:fw + meta["fwords"]])
    out_res = []
    for k, p in enumerate(primes):
        p = int(p)
        cp = all_constants(program, meta, p)
        consts = [cp[ci] for ci in const_slots]
        n_cbank = RES_STEP - dim
        lc = consts[n_cbank:]
        v = [int(v0_res[j][k]) for j in range(m)]
        for s in range(depth):
            C = [0] * 16
            for i in range(min(n_cbank, len(consts))):
                C[i] = consts[i]
            for ax in range(dim):
                sn, sd, tn, td = axes[ax]
                num, den = sn * td + s * tn * sd, sd * td
                C[RES_STEP - 1 - ax] = num % p * pow(den % p, p - 2, p) % p if den != 1 else num % p
            C[RES_STEP] = s
            E = run_res(code, C, lc, p, meta["nslots"], v)
            assert all(E[i] is not None and abs(E[i]) < (1 << 31) for i in range(m)), "state not stored reduced"
            v = E[:m]
        out_res.append([x % p for x in v])
    cf = all_constants(program, meta)
    fconsts = [cf[ci] for ci in const_slots]
    vf = [float(x) for x in v0_f] + [0.0] * (16 - m)
    log_scale = 0.0
    for s in range(depth):
        X = [0.0] * 16
        for ax in range(dim):
            sn, sd, tn, td = axes[ax]
            X[ax] = float(sn * td + s * tn * sd) / float(sd * td)
        X[STEP_SLOT] = float(s)
        E = run_fp(fcode, X, fconsts, dim, meta["fnslots"], vf)
        new = [E[i] for i in range(m)]
        mx = max(abs(x) for x in new)
        if not any(x != x for x in new) and mx > 1e10:
            new = [x / mx for x in new]
            log_scale += math.log(mx)
        vf = new + [0.0] * (16 - m)
    res = [[out_res[k][j] for k in range(len(primes))] for j in range(m)]
    return res, vf[:m], log_scale
