"""r x r CMF walks with rational shifts / trajectories (python/dreams_rns/cmf_walk.py).

At step s, axis i evaluates to shift_i + s * trajectory_i, i.e.
(sn*td + s*tn*sd) / (sd*td); residues use inv(sd*td) mod p. The walk starts
from the identity: P(N) = M(0) * M(1) * ... * M(N-1); p = P[0, r-1],
q = P[r-1, r-1] (cmf_walk.py:1-17, 363-431). The state-vector walk is
v(N) = M(N-1) ... M(0) v(0) from an exact rational v(0) (cmf_walk.py:445-531).
"""

from __future__ import annotations

import itertools
from fractions import Fraction
from typing import Any, Dict, List, Optional, Sequence, Tuple

from .compiler import CmfCompiler, CmfProgram
from .engine import BatchPlan, cmf_spec, parse_rational_list, vec_spec

__all__ = ["run_cmf_walk", "run_cmf_walk_vec", "run_cmf_walk_vec_batch", "parse_rational_list",
           "compile_cmf_spec", "shift_to_axis_offsets"]


# ---------------------------------------------------------------- pFq specs
def _product(factors: Sequence[str]) -> str:
    return " * ".join(factors) if factors else "1"


def _elementary(factors: Sequence[str], k: int) -> str:
    """e_k of the factors as a sum of products (combinations in index order);
    "0" when there are fewer than k factors."""
    terms = [" * ".join(c) for c in itertools.combinations(factors, k)] if k <= len(factors) else []
    return " + ".join(terms) if terms else "0"


def _build_per_axis_companion(p: int, q: int, a_params: Sequence[Fraction],
                              b_params: Sequence[Fraction]) -> Tuple[int, Dict[str, str], List[str]]:
    """r x r companion of a pFq CMF with one axis per parameter
    (cmf_walk.py:32-99): r = max(p, q) + 1, axes x0..x{p-1}, y0..y{q-1},
    factors "(x_i + a_i)" / "(y_j + b_j)". Ones on the sub-diagonal; the last
    column holds prod a-factors in row 0, e_k(b) - e_k(a) with k = r - 1 - row
    in the middle rows, prod b-factors in row r - 1 (for r = 2: rows 0 and 1
    are the two products). Returns (r, {"row,col": expression}, axis names);
    the expression strings use the reference's spelling, since the order in
    which SymPy folds them fixes the FP64 shadow's rounding."""
    r = max(p, q) + 1
    axes = [f"x{i}" for i in range(p)] + [f"y{j}" for j in range(q)]
    fa = [f"({axes[i]} + {a_params[i]})" for i in range(p)]
    fb = [f"({axes[p + j]} + {b_params[j]})" for j in range(q)]
    mat = {f"{i},{j}": ("1" if i == j + 1 else "0") for i in range(r) for j in range(r)}
    last = r - 1
    mat[f"0,{last}"] = _product(fa)
    if r == 2:
        mat[f"1,{last}"] = _product(fb)
    else:
        for row in range(1, last):
            k = last - row
            mat[f"{row},{last}"] = f"({_elementary(fb, k)}) - ({_elementary(fa, k)})"
        mat[f"{last},{last}"] = _product(fb)
    return r, mat, axes


def compile_cmf_spec(spec: Dict, trajectory: Optional[Sequence[int]] = None) -> Optional[CmfProgram]:
    """Compile a pFq spec ({"p", "q", "a_params", "b_params", ...}) to the
    per-axis companion program (cmf_walk.py:104-152): m = max(p, q) + 1,
    dim = p + q, directions = trajectory (default all ones). Entries are
    compiled row-major; "0" entries are skipped, "1" is the integer 1. None
    when an entry is not real (SymPy ``I``)."""
    import sympy as sp

    p, q = int(spec["p"]), int(spec["q"])
    a_params = [Fraction(x) for x in spec["a_params"]]
    b_params = [Fraction(x) for x in spec["b_params"]]
    r, mat, axes = _build_per_axis_companion(p, q, a_params, b_params)
    dim = p + q
    traj = [1] * dim if trajectory is None else list(trajectory)
    if len(traj) != dim:
        raise AssertionError(f"trajectory len {len(traj)} != dim {dim}")
    comp = CmfCompiler(m=r, dim=dim, directions=traj)
    axis_symbols = {name: i for i, name in enumerate(axes)}
    local_ns = {name: sp.Symbol(name) for name in axes}
    for i in range(r):
        for j in range(r):
            text = mat[f"{i},{j}"]
            if text == "0":
                continue
            expr = sp.Integer(1) if text == "1" else sp.sympify(text, locals=local_ns)
            if expr.has(sp.I):
                return None
            comp.compile_matrix_entry(i, j, expr, axis_symbols)
    return comp.build()


def shift_to_axis_offsets(shift: Dict, p: int) -> List[int]:
    """Integer start of every axis from a shift record {"nums", "dens"}:
    max(1, 1 + num) per axis; the denominators are not used
    (cmf_walk.py:157-169)."""
    return [max(1, 1 + int(v)) for v in shift.get("nums", [])]


def run_cmf_walk(program: CmfProgram, depth: int, K: int, shift_vals, trajectory_vals=None) -> Dict[str, Any]:
    """GPU CMF walk with the reference's signature and result dict
    (cmf_walk.py:363-431)."""
    spec = cmf_spec(program, depth, K, shift_vals, trajectory_vals)
    with BatchPlan([spec]) as plan:
        return plan.run().walk_dict(0)


def run_cmf_walk_vec(program: CmfProgram, depth: int, K: int, shift_vals, initial_state, acc_idx: int,
                     const_idx: int, trajectory_vals=None) -> Dict[str, Any]:
    """GPU state-vector walk v(N) = M(N-1) ... M(0) v(0) with the reference's
    signature and result dict (cmf_walk.py:445-531): p = v[acc_idx],
    q = v[const_idx], FP64 shadow rescaled by max |v_f| above 1e10. State
    vectors of up to 16 entries and programs of up to 15 axes."""
    spec = vec_spec(program, depth, K, shift_vals, initial_state, acc_idx, const_idx, trajectory_vals)
    with BatchPlan([spec]) as plan:
        return plan.run().walk_dict(0)


def run_cmf_walk_vec_batch(program: CmfProgram, depth: int, K: int, shift_list, initial_state, acc_idx: int,
                           const_idx: int, trajectory_list=None, crt: bool = False) -> List[Dict[str, Any]]:
    """Many state-vector walks of one program in one launch pair (the sweep
    drivers' loop over shifts x trajectories, odd_zeta_sweep.py:481-494):
    walk i uses shift_list[i] and trajectory_list[i] (default: the program's
    directions). With ``crt``, each dict also holds the centred CRT integers
    ``p_crt`` / ``q_crt`` from the GPU CRT pass."""
    n = len(shift_list)
    trajs = trajectory_list if trajectory_list is not None else [None] * n
    if len(trajs) != n:
        raise ValueError("shift_list and trajectory_list differ in length")
    specs = [vec_spec(program, depth, K, sh, initial_state, acc_idx, const_idx, tr) for sh, tr in zip(shift_list, trajs)]
    if not specs:
        return []
    with BatchPlan(specs, crt=crt) as plan:
        res = plan.run()
        out = []
        for i in range(n):
            d = res.walk_dict(i)
            if crt:
                d["p_crt"], d["q_crt"] = res.crt_values(i)
            out.append(d)
        return out
