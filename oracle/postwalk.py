"""Scalar restatement of the sweep's per-walk post-processing (test
infrastructure only): python/a100_sweep.py:426-505 with the constant-bank
helpers of python/dreams_rns/constants.py:159-207, walk by walk, exactly as
the reference loop runs them. It is the checker of the vectorised
``dreams_rns.postwalk.sweep_hits``.

Pinned by tests/test_postwalk.py: closed-form deltas of known convergents
(22/7 and 355/113 against pi), proximity ordering, the FP64-overflow rule of
the CRT check, and hand-built walks whose hit status is fixed by
construction.
"""

from __future__ import annotations

import math
from typing import Any, Dict, List, Sequence

from .walk import centered, crt_reconstruct


def _proximity(est: float, constants, threshold: float) -> List[Dict[str, Any]]:
    """constants.py:159-179."""
    out = []
    for c in constants:
        d = abs(est - c["value_float"])
        if d < threshold:
            out.append({"name": c["name"], "distance": d})
    out.sort(key=lambda x: x["distance"])
    return out


def _delta_exact(p_big: int, q_big: int, c_mp, dps: int) -> float:
    """constants.py:182-207."""
    import mpmath as mp

    mp.mp.dps = dps
    if q_big == 0:
        return float("-inf")
    err = abs(mp.mpf(p_big) / mp.mpf(q_big) - c_mp)
    if err == 0:
        return float("inf")
    log_err = float(mp.log(err))
    log_q = float(mp.log(abs(mp.mpf(q_big))))
    if log_q <= 0:
        return float("-inf")
    return -(1.0 + log_err / log_q)


def walk_record(res: Dict[str, Any], constants: Sequence[Dict[str, Any]], proximity: float,
                dps: int) -> Dict[str, Any] | None:
    """One result dict (p_residues, q_residues, p_float, q_float, log_scale,
    primes) through a100_sweep.py:426-505; None when skipped or no hit."""
    primes = [int(p) for p in res["primes"]]
    p_big, Mp = crt_reconstruct([int(r) for r in res["p_residues"]], primes)
    q_big, _ = crt_reconstruct([int(r) for r in res["q_residues"]], primes)
    p_big = centered(p_big, Mp)
    q_big = centered(q_big, Mp)
    est = res["p_float"] / res["q_float"] if abs(res["q_float"]) > 1e-300 else float("nan")
    log_scale = res.get("log_scale", 0.0)
    q_float = res["q_float"]
    if not math.isfinite(est):
        return None
    crt_ok = True
    if q_big == 0:
        crt_ok = False
    else:
        try:
            cr = float(p_big) / float(q_big)
            if not math.isfinite(cr):
                crt_ok = False
            elif abs(est) > 1e-10:
                crt_ok = abs(cr - est) / abs(est) < 0.01
            else:
                crt_ok = abs(cr - est) < 0.01
        except Exception:
            crt_ok = False
    prox = _proximity(est, constants, proximity)
    best_delta, best_const = float("-inf"), None
    log_q = log_scale + (math.log(abs(q_float)) if abs(q_float) > 1e-300 else 0)
    if prox:
        cands = [next(c for c in constants if c["name"] == m["name"]) for m in prox]
    else:
        cands = [c for c in constants if abs(est - c["value_float"]) < 1.0]
    for c in cands:
        if crt_ok:
            delta = _delta_exact(p_big, q_big, c["value"], dps)
        else:
            try:
                err = abs(est - c["value_float"])
                delta = -(1.0 + math.log(err) / log_q) if err > 0 and log_q > 0 else float("-inf")
            except Exception:
                delta = float("-inf")
        if delta > best_delta:
            best_delta, best_const = delta, c["name"]
    if not (best_delta > 0 or len(prox) > 0):
        return None
    return {"est_float": est, "best_const": best_const, "best_delta": best_delta,
            "p_bits": Mp.bit_length(), "proximity_matches": [m["name"] for m in prox],
            "delta_positive": best_delta > 0}
