"""Pin the oracle before trusting it (CPU only).

* the numpy restatement and the C restatement reproduce the reference's own
  outputs (golden vectors) bit for bit, residues and FP64 shadow;
* independent checks: exact big-integer / rational walks, a per-prime Python
  interpreter (INV(0) = 0), the FMA-chain shadow, textbook Garner and
  sympy's CRT, closed forms for the prime table.
"""

import math
from fractions import Fraction

import numpy as np
import pytest

from conftest import program_from_json

from oracle import cref, exact
from oracle import walk as ow


def _same_walk(res, out, m=2):
    return (list(map(int, res["p_residues"])) == out["p_residues"]
            and list(map(int, res["q_residues"])) == out["q_residues"]
            and float(res["p_float"]).hex() == out["p_float"]
            and float(res["q_float"]).hex() == out["q_float"]
            and float(res["log_scale"]).hex() == out["log_scale"])


def test_numpy_restatement_matches_reference_pcf(golden):
    for w in golden["run_pcf_walk"]:
        if w["depth"] > 300:
            continue  # the full-depth case is covered by the C restatement below
        p = program_from_json(w["program"])
        assert _same_walk(ow.run_pcf_walk(p, w["a0"], w["depth"], w["K"], w["shift_val"]), w["out"]), w["a"]


def test_numpy_restatement_matches_reference_cmf(golden):
    for c in golden["run_cmf_walk"]:
        p = program_from_json(c["program"])
        shift = [tuple(v) for v in c["shift"]]
        traj = None if c["traj"] is None else [tuple(v) for v in c["traj"]]
        assert _same_walk(ow.run_cmf_walk(p, c["depth"], c["K"], shift, traj), c["out"])


def test_c_restatement_matches_reference(golden):
    for w in golden["run_pcf_walk"]:
        p = program_from_json(w["program"])
        res, Pf, ls = cref.pcf_walk(p, w["a0"], w["depth"], w["out"]["primes"], w["shift_val"])
        r = {"p_residues": res[0, 1], "q_residues": res[1, 1], "p_float": Pf[0, 1], "q_float": Pf[1, 1],
             "log_scale": ls}
        assert _same_walk(r, w["out"]), w["a"]
    for c in golden["run_cmf_walk"]:
        p = program_from_json(c["program"])
        tr = c["traj"] or [[d, 1] for d in c["directions"]]
        axes = [(c["shift"][i][0], c["shift"][i][1], tr[i][0], tr[i][1]) for i in range(c["dim"])]
        res, Pf, ls = cref.walk(p, c["depth"], c["out"]["primes"], axes)
        m = c["m"]
        r = {"p_residues": res[0, m - 1], "q_residues": res[m - 1, m - 1], "p_float": Pf[0, m - 1],
             "q_float": Pf[m - 1, m - 1], "log_scale": ls}
        assert _same_walk(r, c["out"])


def test_exact_integer_walk_pins_pcf_residues(golden):
    # no INV in these programs: the walk is an integer matrix product
    for w in golden["run_pcf_walk"]:
        p = program_from_json(w["program"])
        if any(int(i.op) == 10 for i in p.opcodes) or w["depth"] > 400:
            continue
        init = [[1, w["a0"]], [0, 1]]
        P = exact.exact_walk(p, w["depth"], init, [w["shift_val"]], [1])
        for k, pr in enumerate(w["out"]["primes"]):
            assert exact.reduce_mod(P[0][1], pr) == w["out"]["p_residues"][k]
            assert exact.reduce_mod(P[1][1], pr) == w["out"]["q_residues"][k]


def test_exact_rational_walk_pins_cmf_residues(golden):
    for c in golden["run_cmf_walk"]:
        p = program_from_json(c["program"])
        tr = c["traj"] or [[d, 1] for d in c["directions"]]
        shift = [Fraction(a, b) for a, b in c["shift"]]
        traj = [Fraction(a, b) for a, b in tr]
        m = c["m"]
        try:
            P = exact.exact_walk(p, c["depth"], [[int(i == j) for j in range(m)] for i in range(m)], shift, traj)
        except exact.VanishingDenominator:
            continue  # covered by the per-prime interpreter below
        for k, pr in enumerate(c["out"]["primes"]):
            for val, want in ((P[0][m - 1], c["out"]["p_residues"][k]), (P[m - 1][m - 1], c["out"]["q_residues"][k])):
                r = exact.reduce_mod(val, pr)
                if r is not None:
                    assert r == want


def test_modular_interpreter_pins_inv_zero(golden):
    # INV(0) = 0 (Fermat) at the vanishing denominator of 1/(x - 2) and -1/(n+1)
    for c in golden["run_cmf_walk"]:
        if "1/(x - 2)" not in [e for _, _, e in c["matrix"]]:
            continue
        p = program_from_json(c["program"])
        for k, pr in enumerate(c["out"]["primes"]):
            fn = exact.rational_axis_fn(pr, [tuple(v) for v in c["shift"]], [(d, 1) for d in c["directions"]])
            P = exact.modular_walk(p, c["depth"], pr, [[1, 0], [0, 1]], fn)
            assert P[0][1] == c["out"]["p_residues"][k] and P[1][1] == c["out"]["q_residues"][k]
    for w in golden["run_pcf_walk"]:
        if w["shift_val"] != -3:
            continue
        p = program_from_json(w["program"])
        for k, pr in enumerate(w["out"]["primes"]):
            fn = lambda s, pr=pr: [(w["shift_val"] + s) % pr]
            P = exact.modular_walk(p, w["depth"], pr, [[1, w["a0"]], [0, 1]], fn)
            assert P[0][1] == w["out"]["p_residues"][k]


def test_fma_chain_shadow_pins_blas_order(golden):
    # the reference's numpy @ equals sequential FMA chains on the fixture machine
    for w in golden["run_pcf_walk"][:4]:
        p = program_from_json(w["program"])
        fn = lambda s, w=w: [float(w["shift_val"] + s)]
        Pf, ls = exact.fma_shadow(p, w["depth"], [[1.0, float(w["a0"])], [0.0, 1.0]], fn)
        assert float(Pf[0][1]).hex() == w["out"]["p_float"] and float(Pf[1][1]).hex() == w["out"]["q_float"]
        assert float(ls).hex() == w["out"]["log_scale"]


def test_crt_known_values_and_libraries():
    # values and primes of the reference's C++ CRT test (tests/test_matmul_crt.cpp:152-202)
    primes = [2147483647, 2147483629, 2147483587, 2147483579, 2147483549, 2147483543, 2147483497, 2147483489]
    for v in (0, 1, 12345, 1000000007, 1 << 32, 1 << 40, 1 << 50):
        res = [v % p for p in primes]
        assert ow.crt_reconstruct(res, primes) == (v, math.prod(primes))
        assert exact.garner(res, primes) == (v, math.prod(primes))
    from sympy.ntheory.modular import crt

    rng = np.random.default_rng(3)
    pr = [int(p) for p in ow.generate_rns_primes(32)]
    for _ in range(20):
        res = [int(rng.integers(0, p)) for p in pr]
        x, M = ow.crt_reconstruct(res, pr)
        assert (x, M) == exact.garner(res, pr)
        assert int(crt(pr, res)[0]) == x
        assert ow.centered(x, M) in (x, x - M) and abs(ow.centered(x, M)) <= M // 2


def test_prime_table_closed_forms():
    from dreams_rns.primes import generate_rns_primes

    p32 = ow.generate_rns_primes(32)
    assert p32[0] == 2147483647 and p32[-1] == 2147482921
    assert math.prod(int(p) for p in p32).bit_length() == 992
    assert math.prod(int(p) for p in ow.generate_rns_primes(16)).bit_length() == 496
    assert generate_rns_primes(32).tolist() == p32.tolist()
    assert generate_rns_primes(64).tolist() == ow.generate_rns_primes(64).tolist()
    p128 = generate_rns_primes(128)
    assert p128[63] == 2147482237 and p128[-1] == 2147480849
    assert math.prod(int(p) for p in p128[:64]).bit_length() == 1984
    assert math.prod(int(p) for p in p128).bit_length() == 3968


def test_delta_float_closed_form():
    assert ow.compute_dreams_delta_float(float("nan"), 1.0, 0.0, 1.0) == (-1e10, 0.0)
    assert ow.compute_dreams_delta_float(1.0, 1e-301, 0.0, 1.0) == (-1e10, 0.0)
    d, lq = ow.compute_dreams_delta_float(3.0, 2.0, 10.0, 1.5)
    assert d == 100.0 and lq == 10.0 + math.log(2.0)
    d, lq = ow.compute_dreams_delta_float(3.0, 2.0, -5.0, 1.0)
    assert d == -1e10
    d, lq = ow.compute_dreams_delta_float(3.0, 2.0, 10.0, 1.0)
    assert d == -(1.0 + math.log(0.5) / (10.0 + math.log(2.0)))


def test_delta_exact_special_cases_and_value():
    # runner.py:103-125: q = 0 and |q| <= 1 give -inf, p/q == L gives +inf;
    # a value against a second route (the error from a 300-digit pi, logs in
    # FP64), so a dropped 1 + or a sign or a swapped log fails
    import mpmath as mp

    assert ow.compute_dreams_delta_exact(5, 0, 0.5) == float("-inf")
    assert ow.compute_dreams_delta_exact(1, 2, 0.5) == float("inf")
    assert ow.compute_dreams_delta_exact(3, -1, 0.5) == float("-inf")
    mp.mp.dps = 300
    pi300 = +mp.pi
    err = float(abs(mp.mpf(22) / 7 - pi300))
    want = -(1.0 + math.log(err) / math.log(7.0))
    got = ow.compute_dreams_delta_exact(22, 7, mp.nstr(pi300, 250), dps=200)
    assert abs(got - want) <= 1e-13 * abs(want) and abs(got - 2.42929) < 1e-5
    got = ow.compute_dreams_delta_exact(-355, -113, mp.nstr(pi300, 250), dps=200)
    err = float(abs(mp.mpf(355) / 113 - pi300))
    assert abs(got - (-(1.0 + math.log(err) / math.log(113.0)))) <= 1e-13 * abs(got)


def test_oracle_verify_pcf_equals_reference_record(golden):
    # the reference's own verify_pcf output for the README PCF (depth 2000,
    # K = 32) pins the oracle's verify_pcf: every field
    import dreams_rns as D

    rec = golden["verify_pcf"][0]
    got = ow.verify_pcf(*rec["args"], compile_pcf=D.compile_pcf, pcf_initial_values=D.pcf_initial_values)
    out = rec["out"]
    for key in ("a", "b", "limit", "depth", "K", "p_bits"):
        assert got[key] == out[key], key
    for key in ("target", "est_float", "delta_exact"):
        assert float(got[key]).hex() == out[key], key
    for key in ("delta_float", "log_q"):
        want = float.fromhex(out[key])
        assert abs(got[key] - want) <= 1e-12 * abs(want), key


def test_readme_known_answer_delta_float(golden):
    # README.md:19-32: a = 2, b = n^2, L = 2/(4 - pi), depth 2000, K 32
    ref = golden["verify_pcf"][0]["out"]
    assert abs(float.fromhex(ref["delta_float"]) - (-0.998855)) < 5e-7
    assert ref["p_bits"] == 992


def _zeta_rows():
    import odd_zeta as oz

    fx = oz.load()
    progs = {name: oz.compile_spec(spec) for name, spec in fx["specs"].items()}
    return oz, fx, progs


def test_vec_walk_oracle_matches_reference_odd_zeta_estimates():
    # 126 state-vector walks (zeta(5), zeta(7), zeta(9): 4x4 / 6x6 / 8x8,
    # 6 / 9 / 12 axes, rational shifts and trajectories) at depth 50, K = 16:
    # the reference's p_float / q_float rounded to 15 digits
    oz, fx, progs = _zeta_rows()
    for row in fx["results"]:
        spec = fx["specs"][row["cmf"]]
        v0 = oz.initial_state(spec["n"])
        r = ow.run_cmf_walk_vec(progs[row["cmf"]], row["depth"], row["K"], oz.parse_vec(row["shift"]), v0,
                                spec["accumulator_idx"], spec["constant_idx"], oz.parse_vec(row["trajectory"]))
        est = r["p_float"] / r["q_float"]
        assert abs(est - row["est"]) <= 1e-12 * abs(row["est"]) + 1e-15, (row, est)


def test_vec_walk_oracle_residues_pinned_by_exact_rational_walk():
    # the exact rational state vector reduced mod p equals the RNS residues
    oz, fx, progs = _zeta_rows()
    for row in fx["results"][::21]:
        spec = fx["specs"][row["cmf"]]
        v0 = oz.initial_state(spec["n"])
        shift, traj = oz.parse_vec(row["shift"]), oz.parse_vec(row["trajectory"])
        prog = progs[row["cmf"]]
        r = ow.run_cmf_walk_vec(prog, 12, 8, shift, v0, spec["accumulator_idx"], spec["constant_idx"], traj)
        v = exact.exact_vec_walk(prog, 12, v0, shift, traj)
        for key, idx in (("p_residues", spec["accumulator_idx"]), ("q_residues", spec["constant_idx"])):
            want = [exact.reduce_mod(v[idx], int(p)) for p in r["primes"]]
            assert r[key].tolist() == want, (row, key)


def test_vec_walk_oracle_is_the_transposed_matrix_walk():
    # v(N)^T = v(0)^T M(0)^T ... M(N-1)^T: the matrix walk of the transposed
    # program from P0 = rows v(0)^T gives the same residues and shadow bits
    from dreams_rns.compiler import CmfProgram, Instruction, Opcode

    oz, fx, progs = _zeta_rows()
    spec = fx["specs"]["zeta_5"]
    prog = progs["zeta_5"]
    tops = [Instruction(i.op, i.arg0, i.arg2, i.arg1) if i.op == Opcode.STORE else i for i in prog.opcodes]
    tprog = CmfProgram(prog.m, prog.dim, tops, prog.constants, prog.directions)
    v0 = oz.initial_state(spec["n"])
    shift, traj = oz.parse_vec("[2,3,2,3,1,2]"), oz.parse_vec("[1/3,2,1,2,1,1]")
    r = ow.run_cmf_walk_vec(prog, 30, 8, shift, v0, spec["accumulator_idx"], spec["constant_idx"], traj)
    pp = ow.generate_rns_primes(8)
    m = prog.m
    init_res = np.zeros((m * m, 8), dtype=np.int64)
    for j in range(m):
        init_res[j] = [exact.reduce_mod(v0[j], int(p)) for p in pp]   # row 0 = v0
    init_f = np.zeros(m * m)
    init_f[:m] = [float(x) for x in v0]
    axes = [(s.numerator, s.denominator, t.numerator, t.denominator) for s, t in zip(shift, traj)]
    res, Pf, ls = cref.walk(tprog, 30, pp, axes, init_res, init_f)
    assert res[0, spec["accumulator_idx"]].tolist() == r["p_residues"].tolist()
    assert res[0, spec["constant_idx"]].tolist() == r["q_residues"].tolist()
    assert abs(Pf[0, spec["accumulator_idx"]] / Pf[0, spec["constant_idx"]] - r["p_float"] / r["q_float"]) <= \
        1e-13 * abs(r["p_float"] / r["q_float"])


def test_vec_walk_oracle_matches_reference_exact_ratios_at_depth_5000():
    # the reference's exact big-integer runs (odd_zeta_exact_results.jsonl:
    # zeta(5), depth 5000, k = shift + step * trajectory) pin the oracle's
    # FP64 state-vector walk on the multi-axis spec along the same line
    import odd_zeta as oz

    fx = oz.load()
    spec = fx["specs"]["zeta_5"]
    prog = oz.compile_spec(spec)
    v0 = oz.initial_state(spec["n"])
    picks = [e for e in fx["exact"] if (e["shift"], e["trajectory"]) in ((1, 1), (2, 3), (10, 7))]
    assert len(picks) == 3
    for e in picks:
        sh, tr = oz.exact_walk_axes(spec, e["shift"], e["trajectory"])
        r = ow.run_cmf_walk_vec(prog, e["depth"], 16, sh, v0, spec["accumulator_idx"], spec["constant_idx"], tr)
        est = float(r["p_float"] / r["q_float"])
        assert abs(est - float(e["ratio"])) <= 1e-13 * float(e["ratio"]), e
