"""GPU state-vector walks (run_cmf_walk_vec, cmf_walk.py:445-531) against the
reference's odd-zeta results and the oracle."""

from fractions import Fraction

import numpy as np
import pytest

import dreams_rns as D
import odd_zeta as oz
from dreams_rns.engine import BatchPlan, vec_spec
from oracle import walk as ow

pytestmark = pytest.mark.gpu


def _rows_by_cmf():
    fx = oz.load()
    by = {}
    for r in fx["results"]:
        by.setdefault(r["cmf"], []).append(r)
    return fx, by


def test_odd_zeta_estimates_match_reference_results(gpu_ctx):
    # 126 walks: zeta(5) 4x4 / 6 axes, zeta(7) 6x6 / 9 axes, zeta(9) 8x8 /
    # 12 axes; depth 50, K = 16, rational shifts and trajectories. est =
    # p_float / q_float as the sweep driver writes it (15 digits)
    fx, by = _rows_by_cmf()
    for name, rows in by.items():
        spec = fx["specs"][name]
        prog = oz.compile_spec(spec)
        v0 = oz.initial_state(spec["n"])
        depth, K = rows[0]["depth"], rows[0]["K"]
        assert all(r["depth"] == depth and r["K"] == K for r in rows)
        out = D.run_cmf_walk_vec_batch(prog, depth, K, [oz.parse_vec(r["shift"]) for r in rows], v0,
                                       spec["accumulator_idx"], spec["constant_idx"],
                                       [oz.parse_vec(r["trajectory"]) for r in rows], crt=True)
        print(name, len(rows), "walks", flush=True)
        for i, (r, d) in enumerate(zip(rows, out)):
            est = d["p_float"] / d["q_float"]
            assert abs(est - r["est"]) <= 1e-12 * abs(r["est"]) + 1e-15, (name, r, est)
            if i % 7 == 0:
                ref = ow.run_cmf_walk_vec(prog, depth, K, oz.parse_vec(r["shift"]), v0, spec["accumulator_idx"],
                                          spec["constant_idx"], oz.parse_vec(r["trajectory"]))
                assert d["p_residues"].tolist() == ref["p_residues"].tolist(), (name, r)
                assert d["q_residues"].tolist() == ref["q_residues"].tolist(), (name, r)
                assert abs(d["log_scale"] - ref["log_scale"]) <= 1e-12 * max(1.0, abs(ref["log_scale"]))
                pl = [int(p) for p in ref["primes"]]
                xp, M = ow.crt_reconstruct(ref["p_residues"].tolist(), pl)
                xq, _ = ow.crt_reconstruct(ref["q_residues"].tolist(), pl)
                assert (d["p_crt"], d["q_crt"]) == (ow.centered(xp, M), ow.centered(xq, M)), (name, r)


def test_vec_walk_single_call_and_long_depth(gpu_ctx):
    # the reference's entry point, zeta(5) at depth 1200 (rescales), K = 32
    fx = oz.load()
    spec = fx["specs"]["zeta_5"]
    prog = oz.compile_spec(spec)
    v0 = oz.initial_state(spec["n"])
    shift, traj = [2, 3, 2, 3, 1, 2], [Fraction(2, 3), 2, 1, 2, 1, 1]
    got = D.run_cmf_walk_vec(prog, 1200, 32, shift, v0, spec["accumulator_idx"], spec["constant_idx"], traj)
    ref = ow.run_cmf_walk_vec(prog, 1200, 32, shift, v0, spec["accumulator_idx"], spec["constant_idx"], traj)
    assert got["p_residues"].tolist() == ref["p_residues"].tolist()
    assert got["q_residues"].tolist() == ref["q_residues"].tolist()
    for key in ("p_float", "q_float"):
        assert abs(got[key] - ref[key]) <= 1e-12 * abs(ref[key]), key
    assert abs(got["log_scale"] - ref["log_scale"]) <= 1e-12 * max(1.0, abs(ref["log_scale"]))
    assert got["primes"].tolist() == ref["primes"].tolist()
    # a growing walk: the shadow rescales many times over 1200 steps
    cf = lambda mat, m, dim, axes, dirs: D.compile_cmf_from_dict(mat, m, dim, axes, dirs)
    grow = cf({(0, 0): "x + 2", (0, 1): "x**2", (1, 0): "1", (1, 1): "3*x - 1"}, 2, 1, ["x"], [1])
    g = D.run_cmf_walk_vec(grow, 1200, 32, [1], [Fraction(1), Fraction(1, 2)], 0, 1)
    r = ow.run_cmf_walk_vec(grow, 1200, 32, [1], [Fraction(1), Fraction(1, 2)], 0, 1)
    assert r["log_scale"] > 1000
    assert g["p_residues"].tolist() == r["p_residues"].tolist()
    for key in ("p_float", "q_float"):
        assert abs(g[key] - r[key]) <= 1e-12 * abs(r[key]), key
    assert abs(g["log_scale"] - r["log_scale"]) <= 1e-12 * abs(r["log_scale"])


def test_vec_walk_edge_cases(gpu_ctx):
    # a vanishing denominator on the way (INV(0) = 0 in residues, 0.0 in
    # FP64), a zero row, depth 0 and 1, mixed depths and K in one batch, and
    # a graph-captured plan replayed
    cf = lambda mat, m, dim, axes, dirs: D.compile_cmf_from_dict(mat, m, dim, axes, dirs)
    prog = cf({(0, 0): "1/(x - 3)", (0, 1): "y", (1, 1): "x*y + 1", (2, 0): "1", (2, 2): "2/(y + 1)"},
              3, 2, ["x", "y"], [1, 1])
    v0 = [Fraction(1, 3), Fraction(-2), Fraction(5, 7)]
    cases = [(d, K, sh, tr) for d, K in ((0, 8), (1, 16), (9, 8), (40, 31), (40, 64))
             for sh, tr in (([0, 2], [1, 1]), ([Fraction(1, 2), 1], [Fraction(1, 2), 3]))]
    specs = [vec_spec(prog, d, K, sh, v0, 0, 2, tr) for d, K, sh, tr in cases]
    for use_graph in (False, True):
        with BatchPlan(specs, use_graph=use_graph) as plan:
            for _ in range(2 if use_graph else 1):
                res = plan.run()
                for i, (d, K, sh, tr) in enumerate(cases):
                    ref = ow.run_cmf_walk_vec(prog, d, K, sh, v0, 0, 2, tr)
                    got = res.walk_dict(i)
                    assert got["p_residues"].tolist() == ref["p_residues"].tolist(), (i, use_graph)
                    assert got["q_residues"].tolist() == ref["q_residues"].tolist(), (i, use_graph)
                    for key in ("p_float", "q_float"):
                        a, b = float(got[key]), float(ref[key])
                        assert a == b or abs(a - b) <= 1e-12 * max(abs(a), abs(b)), (i, key, a, b)


def test_vec_walks_refuse_a_mixed_batch(gpu_ctx):
    fx = oz.load()
    spec = fx["specs"]["zeta_5"]
    prog = oz.compile_spec(spec)
    v = vec_spec(prog, 5, 8, [2, 3, 2, 3, 1, 2], oz.initial_state(2), 2, 3)
    pcf = D.compile_pcf("n", "n**2")
    from dreams_rns.engine import pcf_spec
    with pytest.raises(ValueError):
        BatchPlan([v, pcf_spec(pcf, 0, 5, 8)])


def test_odd_zeta_exact_ratios_at_depth_5000(gpu_ctx):
    # the 25 exact big-integer runs of the reference (zeta(5), shifts 1..10,
    # trajectories 1..7, depth 5000, odd_zeta_exact.py:113-206): each is the
    # multi-axis walk along k = shift + step * trajectory; est = p_float /
    # q_float within 1e-12 of the exact ratio, and one walk's residues and
    # CRT equal to the oracle's
    fx = oz.load()
    spec = fx["specs"]["zeta_5"]
    prog = oz.compile_spec(spec)
    v0 = oz.initial_state(spec["n"])
    rows = fx["exact"]
    axes = [oz.exact_walk_axes(spec, e["shift"], e["trajectory"]) for e in rows]
    out = D.run_cmf_walk_vec_batch(prog, 5000, 16, [a[0] for a in axes], v0, spec["accumulator_idx"],
                                   spec["constant_idx"], [a[1] for a in axes], crt=True)
    print("odd zeta exact", len(rows), "walks at depth 5000", flush=True)
    for e, d in zip(rows, out):
        assert e["depth"] == 5000
        est = d["p_float"] / d["q_float"]
        assert abs(est - float(e["ratio"])) <= 1e-12 * float(e["ratio"]), (e, est)
    ref = ow.run_cmf_walk_vec(prog, 5000, 16, axes[7][0], v0, spec["accumulator_idx"], spec["constant_idx"], axes[7][1])
    assert out[7]["p_residues"].tolist() == ref["p_residues"].tolist()
    assert out[7]["q_residues"].tolist() == ref["q_residues"].tolist()
    pl = [int(p) for p in ref["primes"]]
    xp, M = ow.crt_reconstruct(ref["p_residues"].tolist(), pl)
    assert out[7]["p_crt"] == ow.centered(xp, M)
