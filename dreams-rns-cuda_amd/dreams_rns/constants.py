"""Bank of target constants for matching walk limits
(python/dreams_rns/constants.py).

The bank's names, SymPy expressions and descriptions are an interface: the
sweep records name the best constant by its bank name, so the table lists the
same 40 constants as the reference (constants.py:21-67): zeta(2)..zeta(30),
five pi variants, Catalan's G, Euler's gamma, ln 2, e, sqrt 2 and phi.
"""

from __future__ import annotations

import math
import sys
from typing import Any, Dict, List, Tuple

__all__ = ["CONSTANTS_REGISTRY", "load_constants", "compute_match_digits", "match_against_constants",
           "compute_delta_against_constant"]

_ZETA_NOTES = {2: "ζ(2) = π²/6", 3: "Apéry's constant ζ(3)", 4: "ζ(4) = π⁴/90", 6: "ζ(6) = π⁶/945"}

CONSTANTS_REGISTRY: List[Dict[str, str]] = (
    [{"name": f"zeta{k}", "sympy_expr": f"zeta({k})", "description": _ZETA_NOTES.get(k, f"ζ({k})")}
     for k in range(2, 31)]
    + [{"name": n, "sympy_expr": e, "description": d} for n, e, d in (
        ("pi", "pi", "π"),
        ("pi2", "pi**2", "π²"),
        ("pi_sq6", "pi**2/6", "π²/6 = ζ(2)"),
        ("1/pi", "1/pi", "1/π"),
        ("4/pi", "4/pi", "4/π"),
        ("catalan", "catalan", "Catalan's constant G"),
        ("euler_gamma", "EulerGamma", "Euler-Mascheroni γ"),
        ("ln2", "log(2)", "ln(2)"),
        ("e", "E", "Euler's number e"),
        ("sqrt2", "sqrt(2)", "√2"),
        ("phi", "(1+sqrt(5))/2", "Golden ratio φ"),
    )]
)


def load_constants(dps: int = 200) -> List[Dict[str, Any]]:
    """Every bank constant to ``dps`` digits (constants.py:70-106): dicts
    with name, value (mpmath mpf, from SymPy's N at dps + 10 digits),
    value_float, sympy_expr, description. Sets mpmath's working precision
    to ``dps``, as the reference does. A constant that fails to evaluate is
    left out with a warning."""
    if hasattr(sys, "set_int_max_str_digits"):
        sys.set_int_max_str_digits(max(sys.get_int_max_str_digits(), dps * 4))
    import mpmath as mp
    import sympy as sp

    mp.mp.dps = dps
    ns = {"pi": sp.pi, "E": sp.E, "zeta": sp.zeta, "catalan": sp.Catalan, "EulerGamma": sp.EulerGamma,
          "log": sp.log, "sqrt": sp.sqrt}
    out = []
    for entry in CONSTANTS_REGISTRY:
        try:
            val = mp.mpf(str(sp.N(sp.sympify(entry["sympy_expr"], locals=ns), dps + 10)))
        except Exception as e:  # the reference reports and skips
            print(f"WARNING: Could not evaluate constant '{entry['name']}': {e}")
            continue
        out.append({"name": entry["name"], "value": val, "value_float": float(val),
                    "sympy_expr": entry["sympy_expr"], "description": entry["description"]})
    return out


def _digits(err: float, scale: float) -> float:
    if err == 0:
        return 16.0
    return -math.log10(err / max(abs(scale), 1e-300))


def compute_match_digits(estimate: float, constants: List[Dict[str, Any]]) -> Tuple[str, float]:
    """Best constant and its matching digits (constants.py:109-156). Near 1
    (|estimate - 1| < 0.5, against constants also within 0.5 of 1) the digits
    compare the tails estimate - 1 and c - 1, so the zeta values of large
    argument (all close to 1) do not all match; elsewhere they are
    -log10(|estimate - c| / |c|). Returns ("none", 0.0) when nothing scores
    above -1; digits are clipped below at 0."""
    near_one = abs(estimate - 1.0) < 0.5
    best_name, best = "none", -1.0
    for c in constants:
        cv = c["value_float"]
        if near_one and abs(cv - 1.0) < 0.5:
            tail = cv - 1.0
            if abs(tail) < 1e-300:
                continue
            d = _digits(abs((estimate - 1.0) - tail), tail)
        else:
            err = abs(estimate - cv)
            if err == 0:
                d = 16.0
            elif abs(cv) > 1e-300:
                d = _digits(err, cv)
            else:
                d = 0.0
        if d > best:
            best, best_name = d, c["name"]
    return best_name, max(best, 0.0)


def match_against_constants(estimate: float, constants: List[Dict[str, Any]],
                            proximity_threshold: float = 1e-3) -> List[Dict[str, Any]]:
    """Constants within ``proximity_threshold`` of the estimate, nearest
    first (constants.py:159-179): dicts with name, distance, value_float,
    description."""
    hits = [{"name": c["name"], "distance": abs(estimate - c["value_float"]), "value_float": c["value_float"],
             "description": c["description"]}
            for c in constants if abs(estimate - c["value_float"]) < proximity_threshold]
    hits.sort(key=lambda h: h["distance"])
    return hits


def compute_delta_against_constant(p_big: int, q_big: int, constant_mp, dps: int = 200) -> float:
    """delta = -(1 + ln|p/q - c| / ln|q|) at ``dps`` digits
    (constants.py:182-207): -inf for q = 0 or ln|q| <= 0, +inf when p/q
    equals c at that precision. Sets mpmath's working precision."""
    import mpmath as mp

    mp.mp.dps = dps
    if q_big == 0:
        return float("-inf")
    err = abs(mp.mpf(p_big) / mp.mpf(q_big) - constant_mp)
    if err == 0:
        return float("inf")
    log_q = float(mp.log(abs(mp.mpf(q_big))))
    if log_q <= 0:
        return float("-inf")
    return -(1.0 + float(mp.log(err)) / log_q)
