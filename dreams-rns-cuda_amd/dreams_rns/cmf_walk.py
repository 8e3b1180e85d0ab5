"""r x r CMF walks with rational shifts / trajectories (python/dreams_rns/cmf_walk.py).

At step s, axis i evaluates to shift_i + s * trajectory_i, i.e.
(sn*td + s*tn*sd) / (sd*td); residues use inv(sd*td) mod p. The walk starts
from the identity: P(N) = M(0) * M(1) * ... * M(N-1); p = P[0, r-1],
q = P[r-1, r-1] (cmf_walk.py:1-17, 363-431). The state-vector walk is
v(N) = M(N-1) ... M(0) v(0) from an exact rational v(0) (cmf_walk.py:445-531).
"""

from __future__ import annotations

from typing import Any, Dict, List

from .compiler import CmfProgram
from .engine import BatchPlan, cmf_spec, parse_rational_list, vec_spec

__all__ = ["run_cmf_walk", "run_cmf_walk_vec", "run_cmf_walk_vec_batch", "parse_rational_list"]


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
    ``p_crt`` / ``q_crt`` from the GPU Garner pass."""
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
