"""r x r CMF walks with rational shifts / trajectories (python/dreams_rns/cmf_walk.py).

At step s, axis i evaluates to shift_i + s * trajectory_i, i.e.
(sn*td + s*tn*sd) / (sd*td); residues use inv(sd*td) mod p. The walk starts
from the identity: P(N) = M(0) * M(1) * ... * M(N-1); p = P[0, r-1],
q = P[r-1, r-1] (cmf_walk.py:1-17, 363-431).
"""

from __future__ import annotations

from typing import Any, Dict

from .compiler import CmfProgram
from .engine import BatchPlan, cmf_spec, parse_rational_list

__all__ = ["run_cmf_walk", "parse_rational_list"]


def run_cmf_walk(program: CmfProgram, depth: int, K: int, shift_vals, trajectory_vals=None) -> Dict[str, Any]:
    """GPU CMF walk with the reference's signature and result dict
    (cmf_walk.py:363-431)."""
    spec = cmf_spec(program, depth, K, shift_vals, trajectory_vals)
    with BatchPlan([spec]) as plan:
        return plan.run().walk_dict(0)
