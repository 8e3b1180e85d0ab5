"""Independent pins for the walk oracle (test infrastructure only).

These share no code with ``oracle.walk`` and follow the mathematics rather
than the reference's numpy structure:

* ``exact_walk``      the walk product over exact rationals (Python ints /
                      Fractions). Reduced mod p it must equal the RNS
                      residues whenever no INV operand vanishes mod p
                      (runner.py:338-413, cmf_walk.py:363-431 define the
                      product; this computes it without RNS).
* ``exact_vec_walk``  the state-vector walk v <- M v over the rationals.
* ``modular_walk``    a per-prime, pure-Python-int bytecode interpreter.
                      Covers the INV(0) = 0 rule (runner.py:252-256: Fermat
                      ``pow(0, p-2, p) == 0``) that the rational walk cannot.
* ``fma_shadow``      the FP64 shadow with every ``P_f @ M_f`` entry written
                      as the sequential FMA chain acc = a0*b0;
                      acc = fma(ak, bk, acc) (the order OpenBLAS dgemm uses on
                      the machine the fixtures were made on), the ``> 1e10``
                      divide-rescale, and ``math.log``. FMA is exact rational
                      arithmetic rounded once (Fraction -> float is correctly
                      rounded in CPython).
* ``garner``          textbook mixed-radix CRT.
"""

from __future__ import annotations

import math
from fractions import Fraction
from typing import Dict, List, Optional, Sequence, Tuple

NOP, LOAD_X, LOAD_C, LOAD_N, ADD, SUB, MUL, NEG, POW2, POW3, INV, STORE, END = range(13)


class VanishingDenominator(Exception):
    """An INV operand is exactly zero: the rational walk is undefined."""


def _ops(program):
    return [(int(i.op), int(i.arg0), int(i.arg1), int(i.arg2)) for i in program.opcodes]


def _axis_values(program, step: int, shift, traj) -> List[Fraction]:
    """Axis i at step s = shift_i + s * traj_i (cmf_walk.py:206-207)."""
    return [Fraction(s) + step * Fraction(t) for s, t in zip(shift, traj)]


def eval_exact(program, step: int, axes: Sequence[Fraction]) -> List[List[Fraction]]:
    """Interpret the bytecode over the rationals. Unwritten registers read 0,
    matching the zero-initialised register file of the reference."""
    regs: Dict[int, Fraction] = {}
    m = program.m
    M = [[Fraction(0)] * m for _ in range(m)]
    get = lambda r: regs.get(r, Fraction(0))
    for op, a0, a1, a2 in _ops(program):
        if op == END:
            break
        if op == LOAD_X:
            regs[a1] = axes[a0]
        elif op == LOAD_C:
            regs[a1] = Fraction(int(program.constants[a0]))
        elif op == LOAD_N:
            regs[a0] = Fraction(step)
        elif op == ADD:
            regs[a2] = get(a0) + get(a1)
        elif op == SUB:
            regs[a2] = get(a0) - get(a1)
        elif op == MUL:
            regs[a2] = get(a0) * get(a1)
        elif op == NEG:
            regs[a1] = -get(a0)
        elif op == POW2:
            regs[a1] = get(a0) ** 2
        elif op == POW3:
            regs[a1] = get(a0) ** 3
        elif op == INV:
            x = get(a0)
            if x == 0:
                raise VanishingDenominator(step)
            regs[a1] = 1 / x
        elif op == STORE:
            M[a1][a2] = get(a0)
    return M


def _matmul(A, B):
    n = len(A)
    return [[sum(A[i][k] * B[k][j] for k in range(n)) for j in range(n)] for i in range(n)]


def exact_walk(program, depth: int, init, shift, traj) -> List[List[Fraction]]:
    """init * M(0) * M(1) * ... * M(depth-1) over Q, M(step) at axes
    shift + step * traj."""
    P = [[Fraction(v) for v in row] for row in init]
    for step in range(depth):
        P = _matmul(P, eval_exact(program, step, _axis_values(program, step, shift, traj)))
    return P


def exact_vec_walk(program, depth: int, v0, shift, traj) -> List[Fraction]:
    """v(depth) = M(depth-1) ... M(0) v(0) over Q (the state-vector walk,
    cmf_walk.py:445-531, without RNS)."""
    v = [Fraction(x) for x in v0]
    n = len(v)
    for step in range(depth):
        M = eval_exact(program, step, _axis_values(program, step, shift, traj))
        v = [sum((M[i][j] * v[j] for j in range(n)), Fraction(0)) for i in range(n)]
    return v


def reduce_mod(x: Fraction, p: int) -> Optional[int]:
    """x mod p, or None if p divides the denominator."""
    if x.denominator % p == 0:
        return None
    return x.numerator % p * pow(x.denominator, p - 2, p) % p


def modular_walk(program, depth: int, p: int, init_res, axis_fn) -> List[List[int]]:
    """Residue walk for ONE prime with Python ints. ``axis_fn(step)`` gives
    the per-axis residues. INV is Fermat, hence INV(0) = 0."""
    m = program.m
    ops = _ops(program)
    consts = [int(c) % p for c in program.constants]
    P = [[int(v) % p for v in row] for row in init_res]
    for step in range(depth):
        axes = axis_fn(step)
        regs: Dict[int, int] = {}
        get = lambda r: regs.get(r, 0)
        M = [[0] * m for _ in range(m)]
        for op, a0, a1, a2 in ops:
            if op == END:
                break
            if op == LOAD_X:
                regs[a1] = axes[a0]
            elif op == LOAD_C:
                regs[a1] = consts[a0]
            elif op == LOAD_N:
                regs[a0] = step % p
            elif op == ADD:
                regs[a2] = (get(a0) + get(a1)) % p
            elif op == SUB:
                regs[a2] = (get(a0) - get(a1)) % p
            elif op == MUL:
                regs[a2] = get(a0) * get(a1) % p
            elif op == NEG:
                regs[a1] = -get(a0) % p
            elif op == POW2:
                regs[a1] = get(a0) ** 2 % p
            elif op == POW3:
                regs[a1] = get(a0) ** 3 % p
            elif op == INV:
                regs[a1] = pow(get(a0), p - 2, p)
            elif op == STORE:
                M[a1][a2] = get(a0) % p
        P = [[sum(P[i][k] * M[k][j] for k in range(m)) % p for j in range(m)] for i in range(m)]
    return P


def rational_axis_fn(p: int, shift, traj):
    """LOAD_X residues for rational shift/trajectory given as (num, den):
    (sn*td + s*tn*sd) * inv(sd*td) (cmf_walk.py:221-224, 348-358)."""
    parts = []
    for (sn, sd), (tn, td) in zip(shift, traj):
        parts.append((sn * td, tn * sd, pow(sd * td, p - 2, p)))
    return lambda s: [((c0 + s * c1) % p) * c2 % p for c0, c1, c2 in parts]


# ---------------------------------------------------------------- FP64 shadow
def fma(a: float, b: float, c: float) -> float:
    """Correctly rounded a*b + c (exact rational arithmetic, one rounding)."""
    if not (math.isfinite(a) and math.isfinite(b) and math.isfinite(c)):
        return a * b + c
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def fma_matmul(A, B):
    n = len(A)
    C = [[0.0] * n for _ in range(n)]
    for i in range(n):
        for j in range(n):
            acc = A[i][0] * B[0][j]
            for k in range(1, n):
                acc = fma(A[i][k], B[k][j], acc)
            C[i][j] = acc
    return C


def eval_float(program, step: int, axes_f: Sequence[float]):
    """FP64 interpretation; POW2/POW3 are Python ``**`` like the reference
    (runner.py:304-309)."""
    regs: Dict[int, float] = {}
    m = program.m
    M = [[0.0] * m for _ in range(m)]
    get = lambda r: regs.get(r, 0.0)
    for op, a0, a1, a2 in _ops(program):
        if op == END:
            break
        if op == LOAD_X:
            regs[a1] = axes_f[a0]
        elif op == LOAD_C:
            regs[a1] = float(int(program.constants[a0]))
        elif op == LOAD_N:
            regs[a0] = float(step)
        elif op == ADD:
            regs[a2] = get(a0) + get(a1)
        elif op == SUB:
            regs[a2] = get(a0) - get(a1)
        elif op == MUL:
            regs[a2] = get(a0) * get(a1)
        elif op == NEG:
            regs[a1] = -get(a0)
        elif op == POW2:
            regs[a1] = get(a0) ** 2
        elif op == POW3:
            regs[a1] = get(a0) ** 3
        elif op == INV:
            x = get(a0)
            regs[a1] = 1.0 / x if abs(x) > 1e-300 else 0.0
        elif op == STORE:
            M[a1][a2] = get(a0)
    return M


def fma_shadow(program, depth: int, init_f, axis_f_fn) -> Tuple[List[List[float]], float]:
    """Sequential FP64 shadow: P_f <- P_f @ M_f as FMA chains, then if
    max|P_f| > 1e10 divide by it and add log(max) (runner.py:393-398)."""
    P = [list(map(float, row)) for row in init_f]
    log_scale = 0.0
    for step in range(depth):
        P = fma_matmul(P, eval_float(program, step, axis_f_fn(step)))
        mx = max(abs(v) for row in P for v in row)
        if mx > 1e10:
            P = [[v / mx for v in row] for row in P]
            log_scale += math.log(mx)
    return P, log_scale


# ------------------------------------------------------------------------ CRT
def garner(residues: Sequence[int], primes: Sequence[int]) -> Tuple[int, int]:
    """Mixed-radix CRT: digits v_i with x = v0 + v1 p0 + v2 p0 p1 + ...,
    each v_i in [0, p_i). Returns (x, M)."""
    digits: List[int] = []
    for i, (r, p) in enumerate(zip(residues, primes)):
        # value of the partial mixed-radix number mod p
        acc, scale = 0, 1
        for v, q in zip(digits, primes[:i]):
            acc = (acc + v * scale) % p
            scale = scale * q % p
        digits.append((int(r) - acc) * pow(scale, p - 2, p) % p)
    x, M = 0, 1
    for v, p in zip(digits, primes):
        x += v * M
        M *= int(p)
    return x, M
