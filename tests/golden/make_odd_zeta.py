"""Write tests/golden/odd_zeta.json from the reference's odd-zeta data files.

Reads (as data, nothing is imported or run) python/odd_zeta_specs.jsonl (the
CMF specs: matrix entries, axes, default directions / shifts, accumulator and
constant indices) and python/odd_zeta_results/odd_zeta_results.jsonl (126
walks of run_cmf_walk_vec at depth 50, K=16 with their float estimate
p_float / q_float rounded to 15 digits, written by odd_zeta_sweep.py:486-527),
and the exact depth-5000 ratios of odd_zeta_exact_results.jsonl.
"""

import json
import os
import sys

REF = "/root/reference/python"


def main(out=os.path.join(os.path.dirname(os.path.abspath(__file__)), "odd_zeta.json")):
    specs = {}
    with open(os.path.join(REF, "odd_zeta_specs.jsonl")) as f:
        for line in f:
            if line.strip():
                d = json.loads(line)
                specs[d["name"]] = {k: d[k] for k in ("name", "n", "rank", "dim", "matrix", "axis_names",
                                                       "directions", "default_shifts", "accumulator_idx",
                                                       "constant_idx")}
    rows = []
    with open(os.path.join(REF, "odd_zeta_results", "odd_zeta_results.jsonl")) as f:
        for line in f:
            if line.strip():
                d = json.loads(line)
                rows.append({k: d[k] for k in ("cmf", "shift", "trajectory", "depth", "K", "est")})
    exact = []
    with open(os.path.join(REF, "odd_zeta_exact_results", "odd_zeta_exact_results.jsonl")) as f:
        for line in f:
            if line.strip():
                d = json.loads(line)
                exact.append({k: d[k] for k in ("cmf", "n", "rank", "shift", "trajectory", "depth", "ratio")})
    used = sorted({r["cmf"] for r in rows} | {e["cmf"] for e in exact})
    with open(out, "w") as f:
        json.dump({"specs": {k: specs[k] for k in used if k in specs}, "results": rows, "exact": exact}, f,
                  indent=1)
    print("wrote", out, len(rows), "walks,", len(exact), "exact ratios", file=sys.stderr)


if __name__ == "__main__":
    main()
