"""Odd-zeta state-vector walks from the reference's result files (test helper).

The fixture (tests/golden/odd_zeta.json, written by make_odd_zeta.py) holds
the CMF specs and 126 walks of run_cmf_walk_vec at depth 50, K = 16 with
their float estimate p_float / q_float rounded to 15 digits.
"""

from __future__ import annotations

import json
import os
from fractions import Fraction
from math import comb
from typing import Dict, List

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "odd_zeta.json")


def load() -> Dict:
    with open(FIXTURE) as f:
        return json.load(f)


def compile_spec(spec: Dict):
    """The spec's matrix through compile_cmf_from_dict, as the sweep driver
    compiles it (odd_zeta_sweep.py:55-72)."""
    from dreams_rns.compiler import compile_cmf_from_dict

    mat = {}
    for key, expr in spec["matrix"].items():
        r, c = key.split(",")
        mat[(int(r), int(c))] = expr
    return compile_cmf_from_dict(mat, spec["rank"], spec["dim"], spec["axis_names"], spec["directions"])


def initial_state(n: int) -> List[Fraction]:
    """v(0) of the zeta(2n+1) state-vector CMF (the formula of
    odd_zeta_sweep.py:78-116): v[0] = 1/2, v[n] = the k = 1 term below,
    v[n+1] = 1, zero elsewhere. The k-th term is
    (1/2) A_k (sum_j c_j e_j) / k^3 with A_k = (-1)^(k-1) / C(2k, k), e_j the
    j-th elementary symmetric sum of 1/m^2 over m < k, and c_j = 4 (-1)^j /
    k^(2(n-1-j)) for j < n-1, c_(n-1) = 5 (-1)^(n-1)."""

    def term(k: int) -> Fraction:
        e = [Fraction(1)] + [Fraction(0)] * (n - 1)
        for m in range(1, k):
            x = Fraction(1, m * m)
            for j in range(n - 1, 0, -1):
                e[j] += x * e[j - 1]
        c = [Fraction(4 * (-1) ** j, k ** (2 * (n - 1 - j))) for j in range(n - 1)] + [Fraction(5 * (-1) ** (n - 1))]
        a_k = Fraction((-1) ** (k - 1), comb(2 * k, k))
        return Fraction(1, 2) * a_k * sum(cj * ej for cj, ej in zip(c, e)) / k ** 3

    v = [Fraction(0)] * (2 * n)
    v[0] = Fraction(1, 2)
    v[n] = term(1)
    v[n + 1] = Fraction(1)
    return v


def parse_vec(s) -> List[Fraction]:
    """"[2,3,1/3]" -> Fractions."""
    if isinstance(s, (int, float)):
        return [Fraction(s)]
    return [Fraction(t.strip()) for t in s.strip("[]").split(",")]


def exact_walk_axes(spec: Dict, shift: int, trajectory: int):
    """Shift and trajectory vectors of the multi-axis spec that reproduce the
    one-variable walk k = shift + step * trajectory of the reference's exact
    run (odd_zeta_exact.py:113-206 over the k-form matrix of
    odd_zeta_cmf_generator.py:29-176): each axis of the spec stands for a
    linear function a k + b of k (its direction a and its value b + a at
    k = 1 in the spec's default shift), so its shift is a shift + b and its
    trajectory a trajectory."""
    sh, tr = [], []
    for d, s0 in zip(spec["directions"], spec["default_shifts"]):
        b = s0 - d
        sh.append(d * shift + b)
        tr.append(d * trajectory)
    return sh, tr
