"""pFq CMF sweep host logic: the per-axis companion, the spec compiler, the
shift offsets and the sweep's walk enumeration (no GPU)."""

import json
import os
import random
import re
from fractions import Fraction

import pytest

from dreams_rns.cmf_walk import _build_per_axis_companion, compile_cmf_spec, shift_to_axis_offsets
from dreams_rns.sweep import CmfSweep, load_cmf_specs, load_shifts, load_trajectories, sweep_offsets
from oracle import exact

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "sweep_3f2.json")


@pytest.fixture(scope="module")
def sweep_fix():
    with open(FIX) as f:
        return json.load(f)


def _params(spec):
    return [Fraction(x) for x in spec["a_params"]], [Fraction(x) for x in spec["b_params"]]


def test_companion_equals_reference_n_form(sweep_fix):
    # the reference's spec records carry the companion written in one
    # variable n; putting every axis to n must give those strings exactly
    for spec in sweep_fix["specs"]:
        a, b = _params(spec)
        r, mat, axes = _build_per_axis_companion(spec["p"], spec["q"], a, b)
        assert r == spec["rank"]
        assert axes == spec["axis_names"]
        assert set(mat) == set(spec["matrix"])
        for key, text in mat.items():
            assert re.sub(r"\b[xy]\d+\b", "n", text) == spec["matrix"][key], (spec["name"], key)


def test_companion_small_orders():
    # 1F1: r = 2, the two products in the last column
    r, mat, axes = _build_per_axis_companion(1, 1, [Fraction(1, 2)], [Fraction(3)])
    assert (r, axes) == (2, ["x0", "y0"])
    assert mat == {"0,0": "0", "0,1": "(x0 + 1/2)", "1,0": "1", "1,1": "(y0 + 3)"}
    # 2F1: r = 3, middle row e_1(b) - e_1(a)
    r, mat, axes = _build_per_axis_companion(2, 1, [Fraction(1), Fraction(-1, 3)], [Fraction(2)])
    assert r == 3 and axes == ["x0", "x1", "y0"]
    assert mat["0,2"] == "(x0 + 1) * (x1 + -1/3)"
    assert mat["1,2"] == "((y0 + 2)) - ((x0 + 1) + (x1 + -1/3))"
    assert mat["2,2"] == "(y0 + 2)"
    assert [mat["1,0"], mat["2,1"]] == ["1", "1"]
    assert all(mat[f"{i},{j}"] == "0" for i in range(3) for j in range(2) if i != j + 1)
    # 0F2: no a-factors: the empty product is 1, e_k of nothing is 0
    r, mat, _ = _build_per_axis_companion(0, 2, [], [Fraction(1), Fraction(2)])
    assert r == 3 and mat["0,2"] == "1" and mat["2,2"] == "(y0 + 1) * (y1 + 2)"
    assert mat["1,2"] == "((y0 + 1) + (y1 + 2)) - (0)"


def test_compiled_spec_evaluates_to_companion(sweep_fix):
    import sympy as sp

    rng = random.Random(5)
    for spec in sweep_fix["specs"][:12]:
        a, b = _params(spec)
        traj = [rng.randint(-3, 3) for _ in range(spec["dim"])]
        prog = compile_cmf_spec(spec, trajectory=traj)
        r, mat, axes = _build_per_axis_companion(spec["p"], spec["q"], a, b)
        assert (prog.m, prog.dim, prog.directions) == (r, spec["dim"], traj)
        syms = {name: sp.Symbol(name) for name in axes}
        for _ in range(3):
            vals = [Fraction(rng.randint(-40, 40), rng.randint(1, 9)) for _ in axes]
            got = exact.eval_exact(prog, 0, vals)
            subs = {syms[n]: sp.Rational(v.numerator, v.denominator) for n, v in zip(axes, vals)}
            for i in range(r):
                for j in range(r):
                    want = sp.sympify(mat[f"{i},{j}"], locals=syms).subs(subs)
                    assert got[i][j] == Fraction(int(sp.numer(want)), int(sp.denom(want))), (spec["name"], i, j)


def test_compile_spec_default_trajectory_and_length_check():
    spec = {"p": 2, "q": 1, "a_params": ["1/2", "1"], "b_params": ["3/2"]}
    prog = compile_cmf_spec(spec)
    assert prog.directions == [1, 1, 1] and prog.m == 3 and prog.dim == 3
    with pytest.raises(AssertionError):
        compile_cmf_spec(spec, trajectory=[1, 1])


def test_shift_offsets(sweep_fix):
    assert shift_to_axis_offsets({"nums": [1, 3, -1, -1, -3], "dens": [3, 8, 8, 8, 8]}, 3) == [2, 4, 1, 1, 1]
    assert shift_to_axis_offsets({"nums": [0, -2, 5]}, 2) == [1, 1, 6]
    assert shift_to_axis_offsets({}, 3) == []
    for sh in sweep_fix["shifts"]:
        got = shift_to_axis_offsets(sh, 3)
        assert len(got) == len(sh["nums"]) and min(got) >= 1
        assert got == [n + 1 if n >= 0 else 1 for n in sh["nums"]]
    # padded with 1 / cut to the spec's dim
    assert sweep_offsets([{"nums": [2, 0]}], 1, 4) == [[3, 1, 1, 1]]
    assert sweep_offsets([{"nums": [2, 0, 4]}], 1, 2) == [[3, 1]]


def test_loaders_round_trip(tmp_path, sweep_fix):
    sp = tmp_path / "specs.jsonl"
    sp.write_text("\n".join(json.dumps(s) for s in sweep_fix["specs"][:3]) + "\n\n")
    tp = tmp_path / "traj.json"
    tp.write_text(json.dumps({"dim": 5, "count": 2, "trajectories": sweep_fix["trajectories"][:2]}))
    hp = tmp_path / "shifts.json"
    hp.write_text(json.dumps({"dim": 5, "count": 2, "shifts": sweep_fix["shifts"][:2]}))
    assert load_cmf_specs(str(sp)) == sweep_fix["specs"][:3]
    assert load_trajectories(str(tp)) == sweep_fix["trajectories"][:2]
    assert load_shifts(str(hp)) == sweep_fix["shifts"][:2]


def test_sweep_rows_cover_every_shift_once(sweep_fix):
    specs = sweep_fix["specs"][:2]
    trajs = sweep_fix["trajectories"][:3] + [[1, 1]]  # the last one has the wrong length: skipped
    shifts = sweep_fix["shifts"]
    for dedupe in (True, False):
        sw = CmfSweep(specs, trajs, shifts, depth=10, K=8, dedupe=dedupe)
        rows = list(sw._rows())
        assert [(s, t) for s, t, _ in sw.skipped] == [(0, 3), (1, 3)]
        assert sw.total_walks == 2 * 4 * len(shifts)
        seen = {}
        for si, ti, prog, so, hs in rows:
            assert prog.directions == trajs[ti]
            for h in hs:
                assert (si, ti, h) not in seen
                seen[(si, ti, h)] = so
                assert so == shift_to_axis_offsets(shifts[h], 3)
        assert len(seen) == 2 * 3 * len(shifts)
        n_off = len({tuple(shift_to_axis_offsets(h, 3)) for h in shifts})
        assert len(rows) == 2 * 3 * (n_off if dedupe else len(shifts))
