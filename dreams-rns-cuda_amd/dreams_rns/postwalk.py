"""The step after the walk: CRT-overflow detection, delta against a target or
the constant bank, and the sweep's hit selection.

* ``parse_limit``, ``compute_delta``, ``compute_delta_float`` and
  ``crt_overflowed`` follow python/euler2ai_full_verify.py:95-167.
* ``sweep_hits`` applies the CMF sweep's per-walk post-processing
  (python/a100_sweep.py:426-505) to a whole device batch. The reference runs
  it walk by walk in Python with mpmath, which costs far more than the walks
  themselves; here a vectorised screen over walks x constants finds the
  walks that can be hits (a proximity match, or a candidate delta near or
  above 0, or a ratio too close to a constant for FP64 to judge), and only
  those go through the reference's own scalar computation, so every
  reported hit carries exactly the reference's values.
"""

from __future__ import annotations

import math
from typing import Any, Dict, List, Optional, Sequence

import numpy as np

__all__ = ["parse_limit", "compute_delta", "compute_delta_float", "crt_overflowed", "crt_ratio_ok",
           "walk_hit", "sweep_hits"]


def parse_limit(limit_str: str, dps: int = 200):
    """A limit expression ("2/(4 - pi)") as an mpmath value at ``dps``
    digits; None when it is not real or does not parse
    (euler2ai_full_verify.py:95-108)."""
    import mpmath as mp
    import sympy as sp

    mp.mp.dps = dps
    try:
        expr = sp.sympify(limit_str, locals={"pi": sp.pi, "E": sp.E, "EulerGamma": sp.EulerGamma, "I": sp.I,
                                             "sqrt": sp.sqrt, "log": sp.log, "Catalan": sp.Catalan,
                                             "GoldenRatio": sp.GoldenRatio})
        if expr.has(sp.I):
            return None
        return mp.mpf(str(sp.N(expr, dps)))
    except Exception:
        return None


def compute_delta(p_big: int, q_big: int, target_mp, dps: int = 200) -> float:
    """-(1 + ln|p/q - L| / ln|q|) at ``dps`` digits; -inf for q = 0,
    ln|q| = 0 or any failure, +inf for an exact hit
    (euler2ai_full_verify.py:111-128)."""
    import mpmath as mp

    mp.mp.dps = dps
    try:
        p_mp, q_mp = mp.mpf(p_big), mp.mpf(q_big)
        if q_mp == 0:
            return float("-inf")
        err = abs(p_mp / q_mp - target_mp)
        if err == 0:
            return float("inf")
        log_err = float(mp.log(err))
        log_q = float(mp.log(abs(q_mp)))
        if log_q == 0:
            return float("-inf")
        return -(1.0 + log_err / log_q)
    except Exception:
        return float("-inf")


def compute_delta_float(est: float, target_float: float, log_scale: float, q_float: float) -> float:
    """delta from the FP64 shadow when the CRT integers overflowed:
    ln|q| = log_scale + ln|q_float| (log_scale alone when |q_float| <= 1e-300)
    (euler2ai_full_verify.py:131-149)."""
    try:
        err = abs(est - target_float)
        if err == 0:
            return float("inf")
        log_err = math.log(err)
        log_q = log_scale + math.log(abs(q_float)) if abs(q_float) > 1e-300 else log_scale
        if log_q <= 0:
            return float("-inf")
        return -(1.0 + log_err / log_q)
    except Exception:
        return float("-inf")


def crt_overflowed(p_big: int, q_big: int, est_float: float) -> bool:
    """True when the CRT ratio p/q (through mpmath, so integers beyond the
    FP64 range are fine) is not finite or differs from the shadow estimate by
    more than 1% (absolute 0.01 when |est| <= 1e-10)
    (euler2ai_full_verify.py:152-167)."""
    import mpmath as mp

    if q_big == 0:
        return True
    try:
        r = float(mp.mpf(p_big) / mp.mpf(q_big))
        if not math.isfinite(r) or not math.isfinite(est_float):
            return True
        if abs(est_float) > 1e-10:
            return abs(r - est_float) / abs(est_float) > 0.01
        return abs(r - est_float) > 0.01
    except Exception:
        return True


def crt_ratio_ok(p_big: int, q_big: int, est: float) -> bool:
    """The sweep's CRT check (a100_sweep.py:449-462): float(p) / float(q)
    in FP64, so integers whose conversion overflows count as overflowed;
    then within 1% of the estimate (absolute 0.01 when |est| <= 1e-10)."""
    if q_big == 0:
        return False
    try:
        r = float(p_big) / float(q_big)
        if not math.isfinite(r):
            return False
        if abs(est) > 1e-10:
            return abs(r - est) / abs(est) < 0.01
        return abs(r - est) < 0.01
    except Exception:
        return False


def walk_hit(p_big: int, q_big: int, p_float: float, q_float: float, log_scale: float,
             constants: Sequence[Dict[str, Any]], proximity: float = 1e-3, dps: int = 200) -> Optional[Dict[str, Any]]:
    """One walk through the sweep's post-processing (a100_sweep.py:440-505):
    None when the estimate is not finite or the walk is no hit, else
    {est_float, best_const, best_delta, proximity_matches, delta_positive,
    crt_ok}. A hit has a delta above 0 against some candidate or a constant
    within ``proximity``; the candidates are the proximity matches if there
    are any, else every constant within 1 of the estimate; a candidate's
    delta is exact (mpmath, ``dps`` digits) when the CRT ratio agrees with
    the shadow, else from the shadow."""
    from .constants import compute_delta_against_constant, match_against_constants

    est = p_float / q_float if abs(q_float) > 1e-300 else float("nan")
    if not math.isfinite(est):
        return None
    ok = crt_ratio_ok(p_big, q_big, est)
    prox = match_against_constants(est, constants, proximity)
    log_q = log_scale + (math.log(abs(q_float)) if abs(q_float) > 1e-300 else 0)
    if prox:
        by_name = {c["name"]: c for c in constants}
        cands = [by_name[m["name"]] for m in prox]
    else:
        cands = [c for c in constants if abs(est - c["value_float"]) < 1.0]
    best, best_name = float("-inf"), None
    for c in cands:
        if ok:
            delta = compute_delta_against_constant(p_big, q_big, c["value"], dps)
        else:
            try:
                err = abs(est - c["value_float"])
                delta = -(1.0 + math.log(err) / log_q) if err > 0 and log_q > 0 else float("-inf")
            except Exception:
                delta = float("-inf")
        if delta > best:
            best, best_name = delta, c["name"]
    if not (best > 0 or prox):
        return None
    return {"est_float": est, "best_const": best_name, "best_delta": best,
            "proximity_matches": [m["name"] for m in prox], "delta_positive": best > 0, "crt_ok": ok}


def _bit_lengths(limbs: np.ndarray, K: np.ndarray) -> np.ndarray:
    """Bit length of each little-endian u32 limb row (first K limbs)."""
    n, L = limbs.shape
    cols = np.arange(L)[None, :]
    nz = (limbs != 0) & (cols < K[:, None])
    top = np.where(nz.any(axis=1), L - 1 - np.argmax(nz[:, ::-1], axis=1), -1)
    out = np.zeros(n, dtype=np.int64)
    has = top >= 0
    tv = limbs[np.arange(n), np.maximum(top, 0)].astype(np.uint64)
    bl = np.zeros(n, dtype=np.int64)
    v = tv.copy()
    for _ in range(32):  # bit length of the top limb
        more = v != 0
        bl += more
        v >>= np.uint64(1)
    out[has] = 32 * top[has] + bl[has]
    return out


def sweep_hits(batch, specs: Sequence[Dict], trajectories: Sequence[Sequence[int]],
               constants: Sequence[Dict[str, Any]], proximity: float = 1e-3, dps: int = 200,
               margin: float = 1e-3) -> List[Dict[str, Any]]:
    """The sweep's hit records of one ``SweepBatch`` (built with crt=True),
    one per (spec, trajectory, shift) in sweep order, with the fields of
    a100_sweep.py:486-500. Walks the screen clears (finite estimate, no
    proximity match, every candidate delta below -``margin`` with the ratio
    far enough from each constant for FP64) cannot be hits; the rest go
    through ``walk_hit``."""
    res = batch.result
    if res.limbs is None:
        raise ValueError("sweep_hits needs a batch run with crt=True")
    n = len(batch)
    if n == 0:
        return []
    Ks = res.Ks.astype(np.int64)
    pf = res.out["p_float"].astype(np.float64)
    qf = res.out["q_float"].astype(np.float64)
    ls = res.out["log_scale"].astype(np.float64)
    cv = np.array([c["value_float"] for c in constants], dtype=np.float64)
    with np.errstate(all="ignore"):
        aq = np.abs(qf)
        est = np.where(aq > 1e-300, pf / qf, np.nan)
        fin = np.isfinite(est)
        dist = np.abs(est[:, None] - cv[None, :])
        prox_any = (dist < proximity).any(axis=1)
        cand = dist < 1.0
        log_q = ls + np.where(aq > 1e-300, np.log(aq), 0.0)
        d_float = np.where((dist > 0) & (log_q[:, None] > 0), -(1.0 + np.log(dist) / log_q[:, None]), -np.inf)
    # rows whose CRT integers both fit FP64 may take the exact path
    bp = _bit_lengths(res.limbs[:, 0, :], Ks)
    bq = _bit_lengths(res.limbs[:, 1, :], Ks)
    small = (bp <= 1024) & (bq <= 1024) & (bq > 0)
    maybe = fin & (prox_any | (cand & (d_float > -margin)).any(axis=1))
    for i in np.nonzero(fin & small & ~maybe)[0]:
        # the exact path's delta: screen it with the FP64 ratio of the CRT integers
        p_big, q_big = res.crt_values(int(i))
        if not crt_ratio_ok(p_big, q_big, float(est[i])):
            continue
        r = float(p_big) / float(q_big)
        lq = math.log(abs(float(q_big)))
        e = np.abs(r - cv)
        close = e <= 1e-9 * np.maximum(np.abs(cv), 1.0)
        with np.errstate(all="ignore"):
            de = np.where((e > 0) & (lq > 0), -(1.0 + np.log(e) / lq), np.inf if lq > 0 else -np.inf)
        if ((cand[i] & ((de > -margin) | close)).any()):
            maybe[i] = True
    M_bits = {}
    from .primes import generate_rns_primes

    hits = []
    for i in np.nonzero(maybe)[0]:
        p_big, q_big = res.crt_values(int(i))
        h = walk_hit(p_big, q_big, float(pf[i]), float(qf[i]), float(ls[i]), constants, proximity, dps)
        if h is None:
            continue
        K = int(Ks[i])
        if K not in M_bits:
            M_bits[K] = math.prod(int(p) for p in generate_rns_primes(K)).bit_length()
        si, ti = int(batch.spec_idx[i]), int(batch.traj_idx[i])
        spec = specs[si]
        for h_idx in batch.shift_rows[i]:
            hits.append((si, ti, int(h_idx), {
                "cmf_name": spec.get("name", f"CMF_{si}"), "spec_hash": spec.get("spec_hash", ""),
                "rank": spec.get("rank", 2), "traj_idx": ti, "traj": list(trajectories[ti]),
                "shift_idx": int(h_idx), "shift_offsets": list(batch.offsets[i]), "est_float": h["est_float"],
                "best_const": h["best_const"], "best_delta": h["best_delta"], "p_bits": M_bits[K],
                "proximity_matches": h["proximity_matches"], "delta_positive": h["delta_positive"],
            }))
    hits.sort(key=lambda t: t[:3])  # the reference's loop order: spec, trajectory, shift
    return [t[3] for t in hits]
