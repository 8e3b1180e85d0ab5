"""Write tests/golden/reference_vectors.json from the reference implementation.

Run in a container that has the reference checkout (read-only) at
``/root/reference``:

    python tests/golden/make_golden.py [--ref /root/reference/python]

Only the functions whose output mathematics alone cannot pin are exercised:

* ``compile_pcf`` / ``compile_cmf_from_dict``: the exact instruction order,
  which follows SymPy's argument order and fixes the FP64 fold order;
* ``run_pcf_walk`` / ``run_cmf_walk``: the FP64 shadow bits (p_float,
  q_float, log_scale) from numpy/OpenBLAS on this machine; the residues come
  along in the same dicts;
* ``verify_pcf``: the README known-answer case.

The output is data only (inputs and outputs); floats are stored with
``float.hex`` so they round-trip bit for bit.
"""

from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def _load_synth():
    path = os.path.join(REPO, "dreams-rns-cuda_amd", "dreams_rns", "synth.py")
    spec = importlib.util.spec_from_file_location("drx_synth_for_golden", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _prog(p):
    return {
        "m": p.m, "dim": p.dim,
        "ops": [[int(i.op), i.arg0, i.arg1, i.arg2] for i in p.opcodes],
        "constants": [str(int(c)) for c in p.constants],
        "directions": list(p.directions),
    }


def _walk_out(r):
    return {
        "p_residues": [int(v) for v in r["p_residues"]],
        "q_residues": [int(v) for v in r["q_residues"]],
        "p_float": float(r["p_float"]).hex(),
        "q_float": float(r["q_float"]).hex(),
        "log_scale": float(r["log_scale"]).hex(),
        "primes": [int(v) for v in r["primes"]],
    }


def _rat(v):
    if isinstance(v, Fraction):
        return [v.numerator, v.denominator]
    if isinstance(v, tuple):
        return [v[0], v[1]]
    return [int(v), 1]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference/python")
    ap.add_argument("--out", default=os.path.join(HERE, "reference_vectors.json"))
    args = ap.parse_args()
    sys.path.insert(0, args.ref)
    import dreams_rns as ref  # the reference package
    from dreams_rns.cmf_walk import run_cmf_walk
    import numpy as np

    synth = _load_synth()
    out = {"source": "reference python/dreams_rns (runner.py, compiler.py, cmf_walk.py)",
           "numpy": np.__version__}

    # ---------------------------------------------------------------- PCFs
    pcf_cases = [
        ("2", "n**2"), ("3*n + 1", "-2*n**2 + n"),
        ("n**3 - 5", "4*n**6 - n**5 + 3*n**4 - 2*n**3 + n**2 - 7*n + 11"),
        ("n/2 + 3", "n**2/3"), ("-n", "-n**2"), ("5", "1"), ("n**4 + 1", "n**7 - n"),
        ("(n+1)**2", "(2*n+1)*(n-3)"), ("3", "-1/(n+1)"), ("n + I", "n"),
        ("2*n+1", "-n**2*(n+1)**(-2)"), ("n - 7", "n**2 - 9"), ("-4", "n**2 - 9"),
    ]
    c1 = synth.pcf1()
    pcf_cases.append((c1["a"], c1["b"]))
    c2 = synth.pcf2(n_walks=24)
    pcf_cases += [(w["a"], w["b"]) for w in c2["walks"]]
    compiled = []
    for a, b in pcf_cases:
        p = ref.compile_pcf(a, b)
        compiled.append({"a": a, "b": b, "program": None if p is None else _prog(p),
                         "a0": None if p is None else ref.pcf_initial_values(a)})
    out["compile_pcf"] = compiled

    walks = [
        ("2", "n**2", 300, 32, 1),
        (c1["a"], c1["b"], 2000, 32, 1),
        ("n**3 - 5", "4*n**6 - n**5 + 3*n**4 - 2*n**3 + n**2 - 7*n + 11", 150, 16, 1),
        ("n/2 + 3", "n**2/3", 100, 8, 1),
        ("3", "-1/(n+1)", 60, 8, -3),
        ("-4", "n**2 - 9", 40, 8, 0),
        ("n - 7", "n**2 - 9", 120, 12, 2),
        ("2*n+1", "-n**2*(n+1)**(-2)", 80, 8, 1),
    ]
    walks += [(w["a"], w["b"], 200, 32, 1) for w in c2["walks"][:6]]
    pcf_walks = []
    for a, b, depth, K, sv in walks:
        p = ref.compile_pcf(a, b)
        a0 = ref.pcf_initial_values(a)
        r = ref.run_pcf_walk(p, a0, depth, K, sv)
        pcf_walks.append({"a": a, "b": b, "a0": a0, "depth": depth, "K": K, "shift_val": sv,
                          "program": _prog(p), "out": _walk_out(r)})
    out["run_pcf_walk"] = pcf_walks

    # ---------------------------------------------------------------- CMFs
    cmf_cases = [
        {"matrix": {(0, 0): "x + 1", (0, 1): "y*2", (1, 0): "1", (1, 1): "x*y - n"},
         "m": 2, "dim": 2, "axes": ["x", "y"], "directions": [1, 2],
         "depth": 50, "K": 8, "shift": [1, 3], "traj": None},
        {"matrix": {(0, 0): "x/(y+2)", (0, 1): "(x**2 - 3*z)/(x + y + 1)", (0, 2): "z**3 - 1",
                    (1, 0): "1", (1, 1): "-x*y*z", (1, 2): "(2*y - 1)**2",
                    (2, 1): "1", (2, 2): "x + y + z + 1/2"},
         "m": 3, "dim": 3, "axes": ["x", "y", "z"], "directions": [1, -2, 1],
         "depth": 40, "K": 16, "shift": [Fraction(1, 3), 2, (5, 2)],
         "traj": [Fraction(1, 2), (3, 4), 1]},
        {"matrix": {(0, 0): "1/(x - 2)", (0, 1): "x", (1, 0): "1", (1, 1): "x**2 + 1"},
         "m": 2, "dim": 1, "axes": ["x"], "directions": [1],
         "depth": 30, "K": 8, "shift": [0], "traj": None},
    ]
    import numpy as _np
    rng = _np.random.default_rng(4242)
    cf = lambda mat, m, dim, axes, dirs: ref.compile_cmf_from_dict(mat, m, dim, axes, dirs)
    for m, lo, hi, inv, depth, K in ((3, 100, 250, True, 25, 16), (3, 100, 250, True, 20, 8),
                                      (4, 10, 400, False, 20, 8)):
        spec = synth.cmf_walk_spec(rng, m, cf, lo, hi, inv)
        cmf_cases.append({"matrix": spec["matrix"], "m": m, "dim": 3, "axes": spec["axes"],
                          "directions": spec["directions"], "depth": depth, "K": K,
                          "shift": spec["shift"], "traj": None})
    cmf_out = []
    for c in cmf_cases:
        p = ref.compile_cmf_from_dict(c["matrix"], c["m"], c["dim"], c["axes"], c["directions"])
        r = run_cmf_walk(p, c["depth"], c["K"], c["shift"], c["traj"])
        cmf_out.append({
            "matrix": [[k[0], k[1], v] for k, v in c["matrix"].items()],
            "m": c["m"], "dim": c["dim"], "axes": c["axes"], "directions": c["directions"],
            "depth": c["depth"], "K": c["K"], "shift": [_rat(v) for v in c["shift"]],
            "traj": None if c["traj"] is None else [_rat(v) for v in c["traj"]],
            "program": _prog(p), "out": _walk_out(r),
        })
    out["run_cmf_walk"] = cmf_out

    # ------------------------------------------------------------ verify_pcf
    v = ref.verify_pcf("2", "n**2", "2/(4 - pi)", depth=2000, K=32)
    out["verify_pcf"] = [{"args": ["2", "n**2", "2/(4 - pi)", 2000, 32],
                          "out": {k: (float(x).hex() if isinstance(x, float) else x)
                                  for k, x in v.items()}}]

    with open(args.out, "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print("wrote", args.out, os.path.getsize(args.out), "bytes")


if __name__ == "__main__":
    main()
