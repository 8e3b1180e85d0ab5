"""Seeded synthetic workloads for the walk engine.

Every generator uses ``numpy.random.default_rng(seed)`` (PCG64) and returns
plain Python data (polynomial strings, dicts), so the same walks can be fed to
any implementation of the reference API. Shapes follow the five benchmark
workloads of this engine:

=====  ===================================================================
pcf1   one PCF, deg a = 1, deg b = 2, coefficients in [-5, 5], depth 2000,
       K = 32
pcf2   4096 PCFs, deg a ~ U{0..3}, deg b ~ U{0..6}, coefficients in
       [-20, 20], depth 2000, K = 32, per-walk target L ~ U(-5, 5), top 64
cmf3   1024 3x3 CMFs in (x, y, z), entries num/den of degree <= 3 with
       coefficients in [-9, 9], 100-250 ops with >= 1 INV, directions in
       [-2, 2]^3 minus 0, shifts in [0, 5]^3, depth 10000, K = 64
mix4   16384 walks: 50 % 2x2 PCFs, 30 % 3x3, 20 % 4x4 CMFs; depth
       log-uniform in [500, 50000]; K from a bit-growth estimate clamped to
       [16, 128]; 10-400 ops; top 128
serve5 Poisson arrivals of requests drawn from the above generators with
       depth log-uniform in [1000, 100000], m in {2,3,4}, K in {32,64,128}
=====  ===================================================================

Leading polynomial coefficients are redrawn until nonzero. This module has no
package-relative imports on purpose (it is also loaded by file path).
"""

from __future__ import annotations

import itertools
import math
from typing import Dict, List, Optional, Tuple

import numpy as np

SEED_BASE = 1000


def _coef(rng, lo: int, hi: int, nonzero: bool = False) -> int:
    while True:
        c = int(rng.integers(lo, hi + 1))
        if c != 0 or not nonzero:
            return c


def poly_str(coefs: List[int], var: str = "n") -> str:
    """Render sum_k coefs[k] * var**k as a SymPy-parsable string."""
    terms = []
    for k, c in enumerate(coefs):
        if c == 0:
            continue
        if k == 0:
            terms.append(f"({c})")
        elif k == 1:
            terms.append(f"({c})*{var}")
        else:
            terms.append(f"({c})*{var}**{k}")
    return " + ".join(terms) if terms else "0"


def random_poly(rng, deg: int, lo: int, hi: int) -> List[int]:
    coefs = [_coef(rng, lo, hi) for _ in range(deg)]
    coefs.append(_coef(rng, lo, hi, nonzero=True))
    return coefs


def pcf_pair(rng, deg_a: int, deg_b: int, lo: int, hi: int) -> Tuple[str, str]:
    return poly_str(random_poly(rng, deg_a, lo, hi)), poly_str(random_poly(rng, deg_b, lo, hi))


# ----------------------------------------------------------------- pcf1 / pcf2
def pcf1(seed: int = SEED_BASE + 1) -> Dict:
    rng = np.random.default_rng(seed)
    a, b = pcf_pair(rng, 1, 2, -5, 5)
    return {"a": a, "b": b, "depth": 2000, "K": 32}


def pcf2(n_walks: int = 4096, seed: int = SEED_BASE + 2, depth: int = 2000, K: int = 32,
         top_k: int = 64) -> Dict:
    rng = np.random.default_rng(seed)
    walks = []
    for _ in range(n_walks):
        da = int(rng.integers(0, 4))
        db = int(rng.integers(0, 7))
        a, b = pcf_pair(rng, da, db, -20, 20)
        walks.append({"a": a, "b": b})
    targets = rng.uniform(-5.0, 5.0, size=n_walks)
    for w, L in zip(walks, targets):
        w["L"] = float(L)
    return {"walks": walks, "depth": depth, "K": K, "top_k": top_k}


# ------------------------------------------------------------------ CMF shapes
_AXES3 = ("x", "y", "z")


def _monomials(axes, max_deg: int):
    out = []
    for exps in itertools.product(range(max_deg + 1), repeat=len(axes)):
        if sum(exps) <= max_deg:
            out.append(exps)
    return out


def _mpoly_str(rng, axes, max_deg: int, n_terms: int, lo: int, hi: int) -> str:
    mons = _monomials(axes, max_deg)
    pick = rng.choice(len(mons), size=min(n_terms, len(mons)), replace=False)
    terms = []
    for idx in sorted(int(i) for i in pick):
        exps = mons[idx]
        c = _coef(rng, lo, hi, nonzero=True)
        factors = [f"({c})"]
        for name, e in zip(axes, exps):
            if e == 1:
                factors.append(name)
            elif e > 1:
                factors.append(f"{name}**{e}")
        terms.append("*".join(factors))
    return " + ".join(terms)


def cmf_entry(rng, axes, max_deg: int = 3, lo: int = -9, hi: int = 9, p_rational: float = 0.5) -> str:
    num = _mpoly_str(rng, axes, max_deg, int(rng.integers(1, 4)), lo, hi)
    if rng.random() < p_rational:
        den = _mpoly_str(rng, axes, max_deg, int(rng.integers(1, 3)), lo, hi)
        return f"({num})/({den})"
    return num


def cmf_matrix(rng, m: int, axes=_AXES3, max_deg: int = 3, lo: int = -9, hi: int = 9,
               p_rational: float = 0.5) -> Dict[Tuple[int, int], str]:
    return {(i, j): cmf_entry(rng, axes, max_deg, lo, hi, p_rational)
            for i in range(m) for j in range(m)}


def _direction(rng, dim: int) -> List[int]:
    while True:
        d = [int(v) for v in rng.integers(-2, 3, size=dim)]
        if any(d):
            return d


def cmf_walk_spec(rng, m: int, compile_fn, min_ops: int, max_ops: int, need_inv: bool,
                  axes=_AXES3, max_tries: int = 200) -> Dict:
    """Draw a CMF until its compiled program has min_ops..max_ops instructions
    (INV included if ``need_inv``). ``compile_fn(matrix, m, dim, axes, dirs)``
    returns a program with ``.opcodes``."""
    for _ in range(max_tries):
        mat = cmf_matrix(rng, m, axes)
        dirs = _direction(rng, len(axes))
        prog = compile_fn(mat, m, len(axes), list(axes), dirs)
        n_ops = len(prog.opcodes)
        has_inv = any(int(i.op) == 10 for i in prog.opcodes)
        if min_ops <= n_ops <= max_ops and (has_inv or not need_inv):
            shift = [int(v) for v in rng.integers(0, 6, size=len(axes))]
            return {"matrix": mat, "m": m, "axes": list(axes), "directions": dirs,
                    "shift": shift, "program": prog}
    raise RuntimeError("could not draw a CMF in the requested op range")


def cmf3(compile_fn, n_walks: int = 1024, seed: int = SEED_BASE + 3, depth: int = 10000,
         K: int = 64) -> Dict:
    rng = np.random.default_rng(seed)
    walks = [cmf_walk_spec(rng, 3, compile_fn, 100, 250, True) for _ in range(n_walks)]
    return {"walks": walks, "depth": depth, "K": K}


# ------------------------------------------------------------------------ mix4
def estimate_bits(max_deg: int, max_coeff: int, depth: int) -> float:
    """Bit growth of the walk product: max_deg * log2(N!) + N * log2(c + 1)
    with Stirling's log2(N!) (python/euler2ai_full_verify.py:40-92)."""
    if depth > 1:
        log2_fact = depth * math.log2(depth / math.e) + 0.5 * math.log2(2 * math.pi * depth)
    else:
        log2_fact = 1.0
    coeff_bits = math.log2(max_coeff + 1) if max_coeff > 1 else 0.0
    return max(max_deg, 1) * log2_fact + depth * coeff_bits


def k_for(bits: float, lo: int = 16, hi: int = 128) -> int:
    return int(min(hi, max(lo, math.ceil(1.5 * bits / 31))))


def mix4(compile_fn, n_walks: int = 16384, seed: int = SEED_BASE + 4, top_k: int = 128) -> Dict:
    rng = np.random.default_rng(seed)
    walks = []
    for _ in range(n_walks):
        depth = int(round(math.exp(rng.uniform(math.log(500), math.log(50000)))))
        u = rng.random()
        if u < 0.5:
            da, db = int(rng.integers(0, 4)), int(rng.integers(0, 7))
            a, b = pcf_pair(rng, da, db, -20, 20)
            K = k_for(estimate_bits(max(da, db), 20, depth))
            walks.append({"kind": "pcf", "a": a, "b": b, "depth": depth, "K": K,
                          "L": float(rng.uniform(-5, 5))})
        else:
            m = 3 if u < 0.8 else 4
            spec = cmf_walk_spec(rng, m, compile_fn, 10, 400, False)
            K = k_for(estimate_bits(3 * m, 9, depth))
            spec.update({"kind": "cmf", "depth": depth, "K": K, "L": float(rng.uniform(-5, 5))})
            walks.append(spec)
    return {"walks": walks, "top_k": top_k}


# ---------------------------------------------------------------------- serve5
def serve5_arrivals(rate: float, duration_s: float, seed: int = SEED_BASE + 5) -> np.ndarray:
    """Poisson arrival times (seconds) at ``rate`` requests/s."""
    rng = np.random.default_rng(seed + int(rate))
    n = int(rng.poisson(rate * duration_s))
    return np.sort(rng.uniform(0.0, duration_s, size=n))


def serve5_request(rng, compile_fn) -> Dict:
    depth = int(round(math.exp(rng.uniform(math.log(1000), math.log(100000)))))
    K = int(rng.choice([32, 64, 128]))
    m = int(rng.choice([2, 3, 4]))
    if m == 2:
        da, db = int(rng.integers(0, 4)), int(rng.integers(0, 7))
        a, b = pcf_pair(rng, da, db, -20, 20)
        return {"kind": "pcf", "a": a, "b": b, "depth": depth, "K": K,
                "L": float(rng.uniform(-5, 5))}
    spec = cmf_walk_spec(rng, m, compile_fn, 10, 400, False)
    spec.update({"kind": "cmf", "depth": depth, "K": K, "L": float(rng.uniform(-5, 5))})
    return spec


# ------------------------------------------- state-vector walks (odd zeta)
# The zeta(5) state-vector CMF of the reference's sweep drivers
# (python/odd_zeta_specs.jsonl, first line): a 4x4 matrix over 6 axes, the
# state v(0) = (1/2, 0, 1, 1) and p = v[2], q = v[3].
ZETA5 = {
    "matrix": {(0, 0): "-x0/(2*x1)", (1, 0): "-x2/(2*x3*x4**2)", (1, 1): "-x2/(2*x3)",
               (2, 0): "-x0/(x1*x5**5) + 5*x2/(4*x3*x4**2*x5**3)", (2, 1): "5*x2/(4*x3*x5**3)",
               (2, 2): "1", (3, 3): "1"},
    "m": 4, "axes": ["x0", "x1", "x2", "x3", "x4", "x5"], "directions": [1, 2, 1, 2, 1, 1],
    "default_shifts": [2, 3, 2, 3, 1, 2], "acc": 2, "const": 3,
}


def zeta5_sweep(compile_fn, n_shift: int = 64, n_traj: int = 256, depth: int = 1000, K: int = 16) -> Dict:
    """A sweep of the zeta(5) walk over shifts x trajectories, as the
    reference's odd-zeta drivers run it (odd_zeta_sweep.py:481-494; its
    documented sweep is 50 shifts x 1000 trajectories at depth 1000, K = 16):
    shifts are the default with one axis moved by an offset in {+-1/4 .. +-3},
    trajectories the default directions with one or two axes scaled by
    rational multipliers. Deterministic."""
    from fractions import Fraction

    z = ZETA5
    dim = len(z["axes"])
    offs = [Fraction(o) for o in ("1", "-1", "1/2", "-1/2", "1/3", "-1/3", "2", "-2", "1/4", "-1/4", "3", "-3")]
    mults = [Fraction(x) for x in ("1/3", "1/2", "2/3", "3/2", "2", "3", "-1", "-1/2", "-1/3", "4/3", "5/2")]
    shifts = [[Fraction(s) for s in z["default_shifts"]]]
    for o, a in itertools.product(offs, range(dim)):
        if len(shifts) >= n_shift:
            break
        v = list(shifts[0])
        v[a] += o
        shifts.append(v)
    trajs = [[Fraction(d) for d in z["directions"]]]
    for mu, a in itertools.product(mults, range(dim)):
        v = list(trajs[0])
        v[a] *= mu
        trajs.append(v)
    for (m1, m2), (a, b) in itertools.product(itertools.product(mults, mults), itertools.combinations(range(dim), 2)):
        if len(trajs) >= n_traj:
            break
        v = list(trajs[0])
        v[a] *= m1
        v[b] *= m2
        trajs.append(v)
    program = compile_fn(z["matrix"], z["m"], dim, z["axes"], z["directions"])
    v0 = [Fraction(1, 2), Fraction(0), Fraction(1), Fraction(1)]
    return {"program": program, "shifts": shifts[:n_shift], "trajs": trajs[:n_traj], "depth": depth, "K": K,
            "v0": v0, "acc": z["acc"], "const": z["const"]}
