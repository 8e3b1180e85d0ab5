"""PCF walk API with the reference's signatures (python/dreams_rns/runner.py).

Convention (runner.py:1-11): M(n) = [[0, b(n)], [1, a(n)]],
P(N) = A * M(1) * ... * M(N) with A = [[1, a(0)], [0, 1]], p = P[0, m-1],
q = P[1, m-1], delta = -(1 + log|p/q - L| / log|q|).

The walk and the CRT run on the GPU through libdreams_hip; compilation
(SymPy) and the exact delta (mpmath) stay on the host, as in the reference.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import _native as N
from .compiler import CmfCompiler, CmfProgram, Opcode
from .engine import BatchPlan, pcf_spec
from .primes import generate_rns_primes

__all__ = [
    "WalkConfig", "generate_rns_primes", "crt_reconstruct", "centered",
    "compute_dreams_delta_float", "compute_dreams_delta_exact", "compile_pcf",
    "pcf_initial_values", "run_pcf_walk", "verify_pcf",
]


@dataclass
class WalkConfig:
    """Walk configuration record (runner.py:23-28)."""

    K: int = 32
    depth: int = 2000
    delta_threshold: float = -2.0


# ------------------------------------------------------------------------ CRT
_crt_ctx = {}


def _ctx_for_primes(primes) -> N.Context:
    """The default context when ``primes`` is a prefix of its prime table,
    otherwise a cached context holding exactly these primes (any distinct
    odd primes below 2^31)."""
    from .primes import is_prime

    pr = np.asarray([int(p) for p in primes], dtype=np.int64)
    ctx = N.context()
    if len(pr) <= N.MAX_K and all((1 << 30) < int(p) < (1 << 31) for p in pr):
        ctx.ensure_primes(len(pr))
        if np.array_equal(ctx.primes[: len(pr)], pr):
            return ctx
    key = tuple(int(p) for p in pr)
    c = _crt_ctx.get(key)
    if c is None:
        if len(key) > N.MAX_K:
            raise ValueError(f"{len(key)} primes: the engine's CRT takes at most {N.MAX_K}")
        for p in key:
            if not (2 < p < (1 << 31) and is_prime(p)):
                raise ValueError(f"modulus {p} is not an odd prime below 2^31: not an RNS prime of this engine")
        if len(set(key)) != len(key):
            raise ValueError("repeated prime: the moduli are not coprime")
        c = N.Context.__new__(N.Context)
        N.Context.__init__(c, ctx.device, n_primes=0)
        arr = np.ascontiguousarray(pr)
        c._check(c.lib.drx_set_primes(c.h, arr.ctypes.data, len(arr)))
        c.primes = arr
        if len(_crt_ctx) >= 4:
            _crt_ctx.pop(next(iter(_crt_ctx))).close()
        _crt_ctx[key] = c
    return c


def crt_reconstruct(residues, primes) -> Tuple[int, int]:
    """x with x = r_i mod p_i, and M = prod(p_i), over the first
    len(residues) primes (runner.py:57-76): an IndexError when there are
    fewer primes than residues or no residue at all, extra primes unused.
    Like the reference, x starts from the first residue as given, so x is
    in [0, M) for canonical residues and r_0 + ((x - r_0) mod M) in general.

    The CRT runs on the GPU (drx_crt: Garner's mixed radix, or the matrix-core
    CRT for many wide rows); the host only joins the returned limbs into a
    Python int. The moduli must be distinct odd primes below 2^31
    (ValueError otherwise; the reference's Fermat inverse assumes primes).
    """
    res_in = [int(r) for r in residues]
    K = len(res_in)
    if K == 0:
        raise IndexError("empty residue list")
    pr = [int(p) for p in primes]
    if len(pr) < K:
        raise IndexError(f"{K} residues but only {len(pr)} primes")
    pr = pr[:K]
    res = np.array([r % p for r, p in zip(res_in, pr)], dtype=np.int64).reshape(1, K)
    ctx = _ctx_for_primes(pr)
    limbs = np.zeros((1, K), dtype=np.uint32)
    neg = np.zeros(1, dtype=np.uint8)
    ctx._check(ctx.lib.drx_crt(ctx.h, res.ctypes.data, 1, K, K, limbs.ctypes.data, neg.ctypes.data))
    M = math.prod(pr)
    mag = int.from_bytes(limbs.tobytes(), "little")
    x = M - mag if neg[0] else mag
    r0 = res_in[0]
    return r0 + (x - r0) % M, M


def centered(x: int, M: int) -> int:
    """Signed representative: x - M if x > M // 2 (runner.py:79-81)."""
    return x - M if x > M // 2 else x


# ---------------------------------------------------------------------- delta
def compute_dreams_delta_float(p_val: float, q_val: float, log_scale: float, target: float):
    """(delta, log|q|) from the FP64 shadow (runner.py:86-100). The batched
    GPU path evaluates the same expression on the device."""
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
    """Exact delta with mpmath at ``dps`` digits (runner.py:103-125). Sets
    ``mpmath.mp.dps`` globally, as the reference does."""
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


# -------------------------------------------------------------- compilation
def compile_pcf(a_str: str, b_str: str) -> Optional[CmfProgram]:
    """Companion matrix [[0, b(n)], [1, a(n)]] as bytecode; n becomes axis 0
    with direction 1. None if a or b contains I (runner.py:130-166)."""
    import sympy as sp
    from sympy.abc import n

    a_expr = sp.sympify(a_str)
    b_expr = sp.sympify(b_str)
    if a_expr.has(sp.I) or b_expr.has(sp.I):
        return None
    s = sp.Symbol("_s")
    a_expr = a_expr.subs(n, s)
    b_expr = b_expr.subs(n, s)
    comp = CmfCompiler(m=2, dim=1, directions=[1])
    axes = {"_s": 0}
    comp.compile_matrix_entry(0, 1, b_expr, axes)
    comp.compile_matrix_entry(1, 0, sp.Integer(1), axes)
    comp.compile_matrix_entry(1, 1, a_expr, axes)
    return comp.build()


def pcf_initial_values(a_str: str) -> int:
    """a(0) as an int (runner.py:169-177)."""
    import sympy as sp
    from sympy.abc import n

    return int(sp.sympify(a_str).subs(n, 0))


# ----------------------------------------------------------------------- walk
def run_pcf_walk(program: CmfProgram, a0: int, depth: int, K: int = 32, shift_val: int = 1):
    """GPU PCF walk with the reference's signature and result dict
    (runner.py:338-413): p/q residues (K,) int64, p/q floats, log_scale,
    primes."""
    assert program.m == 2, "run_pcf_walk only supports 2x2 companion matrices"
    spec = pcf_spec(program, a0, depth, K, shift_val)
    with BatchPlan([spec]) as plan:
        return plan.run().walk_dict(0)


def verify_pcf(a_str: str, b_str: str, limit_str: str, depth: int = 2000, K: int = 32, dps: int = 200):
    """compile -> GPU walk -> GPU CRT -> deltas (runner.py:416-476)."""
    import mpmath as mp
    import sympy as sp

    mp.mp.dps = dps
    program = compile_pcf(a_str, b_str)
    if program is None:
        return None
    a0 = pcf_initial_values(a_str)
    target_expr = sp.sympify(limit_str, locals={"pi": sp.pi, "E": sp.E})
    target_mp = mp.mpf(str(sp.N(target_expr, dps)))
    spec = pcf_spec(program, a0, depth, K, 1, target=float(target_mp))
    with BatchPlan([spec], crt=True) as plan:
        res = plan.run()
    r = res.walk_dict(0)
    delta_float, log_q = compute_dreams_delta_float(r["p_float"], r["q_float"], r["log_scale"],
                                                    float(target_mp))
    p_big, q_big = res.crt_values(0)
    Mp = math.prod(int(p) for p in r["primes"])
    delta_exact = compute_dreams_delta_exact(p_big, q_big, target_mp, dps)
    est = r["p_float"] / r["q_float"] if abs(r["q_float"]) > 1e-300 else float("nan")
    return {
        "a": a_str, "b": b_str, "limit": limit_str, "target": float(target_mp),
        "est_float": est, "delta_float": delta_float, "delta_exact": delta_exact,
        "log_q": log_q, "depth": depth, "K": K, "p_bits": Mp.bit_length(),
    }
