"""numpy restatement of the reference CPU walk (test infrastructure only).

Each function follows the reference function cited in its docstring and keeps
its cost structure (numpy over the K primes per instruction, a fresh 512-entry
register file every step, INV as a Python loop of ``pow``), because this module
is also the timed CPU baseline.

Pinned by (see tests/test_oracle.py):
* golden vectors written by the reference itself (tests/golden/*.json);
* the exact big-integer walk and the pure-Python modular interpreter in
  ``oracle.exact`` (residues, every config shape);
* CRT uniqueness and round trips, sympy's ``crt``;
* the FMA-chain shadow in ``oracle.exact`` (p_float/q_float/log_scale).
"""

from __future__ import annotations

import math
from typing import Any, Dict, List, Sequence, Tuple

import numpy as np

# Opcode numbers of the bytecode ISA (python/dreams_rns/compiler.py:14-28).
NOP, LOAD_X, LOAD_C, LOAD_N, ADD, SUB, MUL, NEG, POW2, POW3, INV, STORE, END = range(13)
NREGS = 512


def generate_rns_primes(K: int) -> np.ndarray:
    """First K primes below 2^31 scanning odd candidates downwards, never
    below 2^30 (runner.py:33-52). Fewer are returned if the range runs out."""
    out: List[int] = []
    c = (1 << 31) - 1
    while len(out) < K and c >= (1 << 30):
        if _is_prime_td(c):
            out.append(c)
        c -= 2
    return np.array(out, dtype=np.int64)


def _is_prime_td(n: int) -> bool:
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    d = 3
    while d * d <= n:
        if n % d == 0:
            return False
        d += 2
    return True


def _ops(program) -> List[Tuple[int, int, int, int]]:
    return [(int(i.op), int(i.arg0), int(i.arg1), int(i.arg2)) for i in program.opcodes]


def const_residues(program, primes: np.ndarray) -> np.ndarray:
    """(n_const, K) table of canonical c mod p (runner.py:182-195)."""
    plist = [int(p) for p in primes]
    tab = np.zeros((len(program.constants), len(plist)), dtype=np.int64)
    for i, c in enumerate(program.constants):
        c = int(c)
        tab[i] = [c % p for p in plist]  # Python % is already canonical
    return tab


def eval_rns(ops, m: int, step: int, axis_res: Sequence[np.ndarray], pp: np.ndarray,
             ctab: np.ndarray) -> np.ndarray:
    """M(step) mod every prime, shape (m, m, K) (runner.py:198-260,
    cmf_walk.py:193-258). ``axis_res[a]`` is the LOAD_X value vector."""
    K = len(pp)
    regs = [np.zeros(K, dtype=np.int64) for _ in range(NREGS)]
    M = np.zeros((m, m, K), dtype=np.int64)
    for op, a0, a1, a2 in ops:
        if op == END:
            break
        if op == LOAD_X:
            regs[a1] = axis_res[a0]
        elif op == LOAD_C:
            regs[a1] = ctab[a0].copy()
        elif op == LOAD_N:
            regs[a0] = np.int64(step) % pp
        elif op == ADD:
            regs[a2] = (regs[a0] + regs[a1]) % pp
        elif op == SUB:
            regs[a2] = (regs[a0] - regs[a1]) % pp
        elif op == MUL:
            regs[a2] = (regs[a0] * regs[a1]) % pp
        elif op == NEG:
            regs[a1] = (pp - regs[a0]) % pp
        elif op == POW2:
            regs[a1] = (regs[a0] * regs[a0]) % pp
        elif op == POW3:
            regs[a1] = (regs[a0] * regs[a0] % pp * regs[a0]) % pp
        elif op == INV:
            regs[a1] = np.array([pow(int(v), int(p) - 2, int(p)) for v, p in zip(regs[a0], pp)],
                                dtype=np.int64)
        elif op == STORE:
            M[a1, a2] = regs[a0] % pp
    return M


def eval_float(ops, m: int, step: int, axis_f: Sequence[float], consts: Sequence[int]) -> np.ndarray:
    """M(step) in FP64 (runner.py:263-316, cmf_walk.py:261-313)."""
    regs = [0.0] * NREGS
    M = np.zeros((m, m), dtype=np.float64)
    for op, a0, a1, a2 in ops:
        if op == END:
            break
        if op == LOAD_X:
            regs[a1] = axis_f[a0]
        elif op == LOAD_C:
            regs[a1] = float(consts[a0])
        elif op == LOAD_N:
            regs[a0] = float(step)
        elif op == ADD:
            regs[a2] = regs[a0] + regs[a1]
        elif op == SUB:
            regs[a2] = regs[a0] - regs[a1]
        elif op == MUL:
            regs[a2] = regs[a0] * regs[a1]
        elif op == NEG:
            regs[a1] = -regs[a0]
        elif op == POW2:
            regs[a1] = regs[a0] ** 2
        elif op == POW3:
            regs[a1] = regs[a0] ** 3
        elif op == INV:
            x = regs[a0]
            regs[a1] = 1.0 / x if abs(x) > 1e-300 else 0.0
        elif op == STORE:
            M[a1, a2] = regs[a0]
    return M


def matmul_mod(A: np.ndarray, B: np.ndarray, pp: np.ndarray) -> np.ndarray:
    """A @ B mod p over the K axis. 2x2 sums both products before the
    ``%`` (runner.py:321-333); r x r reduces after every term
    (cmf_walk.py:174-188). Both give the same canonical residues."""
    r = A.shape[0]
    C = np.zeros_like(A)
    for i in range(r):
        for j in range(r):
            if r == 2:
                C[i, j] = (A[i, 0] * B[0, j] + A[i, 1] * B[1, j]) % pp
            else:
                acc = np.zeros_like(pp)
                for k in range(r):
                    acc = (acc + A[i, k] * B[k, j]) % pp
                C[i, j] = acc
    return C


def _shadow_step(P_f: np.ndarray, M_f: np.ndarray, log_scale: float) -> Tuple[np.ndarray, float]:
    P_f = P_f @ M_f
    mx = np.max(np.abs(P_f))
    if mx > 1e10:
        P_f /= mx
        log_scale += math.log(mx)
    return P_f, log_scale


def run_pcf_walk(program, a0: int, depth: int, K: int = 32, shift_val: int = 1) -> Dict[str, Any]:
    """PCF walk P = A * M(shift) * ... * M(shift+depth-1), A = [[1,a0],[0,1]]
    (runner.py:338-413)."""
    pp = generate_rns_primes(K)
    m = program.m
    assert m == 2, "run_pcf_walk only supports 2x2 companion matrices"
    ops = _ops(program)
    ctab = const_residues(program, pp)
    P = np.zeros((2, 2, len(pp)), dtype=np.int64)
    P[0, 0] = 1
    P[0, 1] = [int(a0) % int(p) for p in pp]
    P[1, 1] = 1
    P_f = np.array([[1.0, float(a0)], [0.0, 1.0]])
    log_scale = 0.0
    dirs = list(program.directions)
    for step in range(depth):
        axis_res = [np.int64(shift_val + step * d) % pp for d in dirs]
        axis_f = [float(shift_val + step * d) for d in dirs]
        M = eval_rns(ops, m, step, axis_res, pp, ctab)
        M_f = eval_float(ops, m, step, axis_f, program.constants)
        P = matmul_mod(P, M, pp)
        P_f, log_scale = _shadow_step(P_f, M_f, log_scale)
    return {
        "p_residues": P[0, 1], "q_residues": P[1, 1],
        "p_float": P_f[0, 1], "q_float": P_f[1, 1],
        "log_scale": log_scale, "primes": pp,
    }


def parse_rationals(vals) -> Tuple[List[int], List[int]]:
    """int / Fraction / (num, den) -> numerators, denominators
    (cmf_walk.py:332-345)."""
    from fractions import Fraction

    nums, dens = [], []
    for v in vals:
        if isinstance(v, Fraction):
            nums.append(v.numerator)
            dens.append(v.denominator)
        elif isinstance(v, tuple):
            nums.append(v[0])
            dens.append(v[1])
        else:
            nums.append(int(v))
            dens.append(1)
    return nums, dens


def run_cmf_walk(program, depth: int, K: int, shift_vals, trajectory_vals=None) -> Dict[str, Any]:
    """r x r CMF walk from the identity with rational shift / trajectory;
    axis i at step s is (sn*td + s*tn*sd) / (sd*td) (cmf_walk.py:363-431)."""
    r = program.m
    pp = generate_rns_primes(K)
    ops = _ops(program)
    ctab = const_residues(program, pp)
    sn, sd = parse_rationals(shift_vals)
    if trajectory_vals is not None:
        tn, td = parse_rationals(trajectory_vals)
    else:
        tn = list(program.directions)
        td = [1] * len(tn)
    inv_cd = []
    for s, t in zip(sd, td):
        cd = s * t
        inv_cd.append(np.ones(len(pp), dtype=np.int64) if cd == 1 else
                      np.array([pow(cd, int(p) - 2, int(p)) for p in pp], dtype=np.int64))
    P = np.zeros((r, r, len(pp)), dtype=np.int64)
    for i in range(r):
        P[i, i] = 1
    P_f = np.eye(r)
    log_scale = 0.0
    for step in range(depth):
        axis_res, axis_f = [], []
        for a in range(len(tn)):
            k_num = sn[a] * td[a] + step * tn[a] * sd[a]
            axis_res.append((np.int64(k_num) % pp * inv_cd[a]) % pp)
            axis_f.append(float(k_num) / float(sd[a] * td[a]))
        M = eval_rns(ops, r, step, axis_res, pp, ctab)
        M_f = eval_float(ops, r, step, axis_f, program.constants)
        P = matmul_mod(P, M, pp)
        P_f, log_scale = _shadow_step(P_f, M_f, log_scale)
    return {
        "p_residues": P[0, r - 1], "q_residues": P[r - 1, r - 1],
        "p_float": P_f[0, r - 1], "q_float": P_f[r - 1, r - 1],
        "log_scale": log_scale, "primes": pp,
    }


def run_cmf_walk_vec(program, depth: int, K: int, shift_vals, initial_state, acc_idx: int, const_idx: int,
                     trajectory_vals=None) -> Dict[str, Any]:
    """State-vector walk v(N) = M(N-1) ... M(0) v(0) with rational shift /
    trajectory and an exact rational initial state (cmf_walk.py:445-531):
    residues of v(0) are num * den^(p-2) mod p; per step v <- M v mod p with a
    reduction after every term (cmf_walk.py:435-441); the shadow is
    v_f <- M_f @ v_f, rescaled by max|v_f| when above 1e10."""
    r = program.m
    pp = generate_rns_primes(K)
    ops = _ops(program)
    ctab = const_residues(program, pp)
    sn, sd = parse_rationals(shift_vals)
    if trajectory_vals is not None:
        tn, td = parse_rationals(trajectory_vals)
    else:
        tn = list(program.directions)
        td = [1] * len(tn)
    inv_cd = []
    for s_, t_ in zip(sd, td):
        cd = s_ * t_
        inv_cd.append(np.ones(len(pp), dtype=np.int64) if cd == 1 else
                      np.array([pow(cd, int(p) - 2, int(p)) for p in pp], dtype=np.int64))
    v = np.zeros((r, len(pp)), dtype=np.int64)
    for i in range(r):
        f = initial_state[i]
        if f == 0:
            continue
        for ki, p in enumerate(int(q) for q in pp):
            num = f.numerator % p
            v[i, ki] = num if f.denominator == 1 else num * pow(f.denominator, p - 2, p) % p
    v_f = np.array([float(f) for f in initial_state], dtype=np.float64)
    log_scale = 0.0
    for step in range(depth):
        axis_res, axis_f = [], []
        for a in range(len(tn)):
            k_num = sn[a] * td[a] + step * tn[a] * sd[a]
            axis_res.append((np.int64(k_num) % pp * inv_cd[a]) % pp)
            axis_f.append(float(k_num) / float(sd[a] * td[a]))
        M = eval_rns(ops, r, step, axis_res, pp, ctab)
        M_f = eval_float(ops, r, step, axis_f, program.constants)
        w = np.zeros_like(v)
        for i in range(r):
            for j in range(r):
                w[i] = (w[i] + M[i, j] * v[j]) % pp
        v = w
        v_f = M_f @ v_f
        mx = np.max(np.abs(v_f))
        if mx > 1e10:
            v_f /= mx
            log_scale += math.log(mx)
    return {
        "p_residues": v[acc_idx], "q_residues": v[const_idx],
        "p_float": v_f[acc_idx], "q_float": v_f[const_idx],
        "log_scale": log_scale, "primes": pp,
    }


def crt_reconstruct(residues, primes) -> Tuple[int, int]:
    """Incremental CRT: x in [0, M), M = prod(primes) (runner.py:57-76)."""
    x = int(residues[0])
    M = int(primes[0])
    for a, p in zip(residues[1:], primes[1:]):
        a, p = int(a), int(p)
        t = ((a - x % p) % p) * pow(M % p, p - 2, p) % p
        x += M * t
        M *= p
    return x, M


def centered(x: int, M: int) -> int:
    """Signed representative (runner.py:79-81)."""
    return x - M if x > M // 2 else x


def compute_dreams_delta_float(p_val: float, q_val: float, log_scale: float, target: float):
    """(delta, log|q|) from the FP64 shadow (runner.py:86-100)."""
    if not math.isfinite(p_val) or not math.isfinite(q_val) or abs(q_val) < 1e-300:
        return -1e10, 0.0
    err = abs(p_val / q_val - target)
    log_q = log_scale + math.log(abs(q_val))
    if err < 1e-300:
        return 100.0, log_q
    if log_q <= 0.0:
        return -1e10, log_q
    return -(1.0 + math.log(err) / log_q), log_q


def compute_dreams_delta_exact(p_big: int, q_big: int, target, dps: int = 200) -> float:
    """mpmath delta of exact p/q (runner.py:103-125); sets mp.dps globally
    like the reference does."""
    import mpmath as mp

    mp.mp.dps = dps
    if q_big == 0:
        return float("-inf")
    err = abs(mp.mpf(p_big) / mp.mpf(q_big) - mp.mpf(str(target)))
    qq = abs(mp.mpf(q_big))
    if err == 0:
        return float("inf")
    if qq <= 1:
        return float("-inf")
    return float(-(1 + mp.log(err) / mp.log(qq)))


def verify_pcf(a_str: str, b_str: str, limit_str: str, depth: int = 2000, K: int = 32,
               dps: int = 200, compile_pcf=None, pcf_initial_values=None):
    """compile -> walk -> CRT -> deltas (runner.py:416-476). The compiler is
    injected so the oracle does not depend on the product package."""
    import mpmath as mp
    import sympy as sp

    mp.mp.dps = dps
    program = compile_pcf(a_str, b_str)
    if program is None:
        return None
    a0 = pcf_initial_values(a_str)
    res = run_pcf_walk(program, a0, depth, K)
    target_mp = mp.mpf(str(sp.N(sp.sympify(limit_str, locals={"pi": sp.pi, "E": sp.E}), dps)))
    delta_float, log_q = compute_dreams_delta_float(res["p_float"], res["q_float"],
                                                    res["log_scale"], float(target_mp))
    primes = [int(p) for p in res["primes"]]
    p_big, Mp = crt_reconstruct([int(r) for r in res["p_residues"]], primes)
    q_big, _ = crt_reconstruct([int(r) for r in res["q_residues"]], primes)
    p_big, q_big = centered(p_big, Mp), centered(q_big, Mp)
    delta_exact = compute_dreams_delta_exact(p_big, q_big, target_mp, dps)
    est = res["p_float"] / res["q_float"] if abs(res["q_float"]) > 1e-300 else float("nan")
    return {
        "a": a_str, "b": b_str, "limit": limit_str, "target": float(target_mp),
        "est_float": est, "delta_float": delta_float, "delta_exact": delta_exact,
        "log_q": log_q, "depth": depth, "K": K, "p_bits": Mp.bit_length(),
    }
