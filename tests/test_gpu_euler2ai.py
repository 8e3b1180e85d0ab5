"""Euler2AI batch verification on the GPU: every row equals the one-PCF-at-a-
time computation of euler2ai_full_verify.py on the oracle's walk and CRT;
the README PCF's delta equals the reference's own verify_pcf golden value."""

import math

import pytest

from dreams_rns.euler2ai import verify_pcf_records, write_csv
from dreams_rns.postwalk import compute_delta, compute_delta_float, crt_overflowed, parse_limit
from dreams_rns.runner import compile_pcf, pcf_initial_values
from oracle import walk as ow

pytestmark = pytest.mark.gpu


def _expected(a, b, limit, depth, K, dps=200):
    prog = compile_pcf(a, b)
    res = ow.run_pcf_walk(prog, pcf_initial_values(a), depth, K)
    primes = [int(p) for p in res["primes"]]
    p_big, M = ow.crt_reconstruct([int(r) for r in res["p_residues"]], primes)
    q_big, _ = ow.crt_reconstruct([int(r) for r in res["q_residues"]], primes)
    p_big, q_big = ow.centered(p_big, M), ow.centered(q_big, M)
    est = res["p_float"] / res["q_float"] if abs(res["q_float"]) > 1e-300 else float("nan")
    target = parse_limit(limit, dps) if limit else None
    if target is None:
        return est, float("nan"), "none", M.bit_length()
    if crt_overflowed(p_big, q_big, est):
        return est, compute_delta_float(est, float(target), res["log_scale"], res["q_float"]), "float", M.bit_length()
    return est, compute_delta(p_big, q_big, target, dps), "crt", M.bit_length()


def test_batch_verification_rows(gpu_ctx, golden, tmp_path):
    recs = [
        {"a": "2", "b": "n**2", "limit": "2/(4 - pi)", "delta": -1.0, "convergence_rate": 0.3,
         "sources": [[[1, 1], [0, 0]], [[2, 1], [1, 0]]]},
        {"a": "3*n + 1", "b": "-2*n**2", "limit": "1/log(2)", "delta": 0.2},
        {"a": "n", "b": "sqrt(-1)*n", "limit": "1"},          # imaginary: SKIP
        {"a": "5*n**2 + 3", "b": "-n**4", "limit": "", "delta": 1.0},
        {"a": "2", "b": "n**2", "limit": "2/(4 - pi)", "delta": 5.0},  # repeated PCF: cached result
        {"a": "(", "b": "n", "limit": "1"},                  # does not compile: error, no row
        {"a": "n + 1", "b": "n", "limit": "E - 1", "delta": 0.0},
    ]
    depth = 300
    for K in (32, 0):
        rows, summ = verify_pcf_records(recs, depth=depth, K=K, dps=200)
        print("euler2ai K", K, summ, flush=True)
        assert [r["task_idx"] for r in rows] == [0, 1, 2, 3, 4, 5, 7]
        assert rows[3]["match"] == "SKIP" and summ["skipped"] == 1 and summ["errors"] == 1
        for r in rows:
            if r["match"] == "SKIP":
                continue
            Kr = r["K"]
            assert Kr == (K if K else max(64, Kr))
            est, delta, method, p_bits = _expected(r["a"], r["b"], r["limit"], depth, Kr)
            assert r["est_float"] == (f"{est:.10f}" if math.isfinite(est) else str(est))
            assert r["delta_cuda_rns"] == (f"{delta:.5f}" if math.isfinite(delta) else str(delta)), r
            assert r["delta_method"] == method and r["p_bits"] == p_bits
        # the repeated PCF reuses the first record's result, delta_ref included
        assert rows[5]["delta_cuda_rns"] == rows[0]["delta_cuda_rns"] and rows[5]["delta_diff"] == rows[0]["delta_diff"]
        assert rows[5]["delta_ref_cpu"] == "5.00000" and rows[4]["delta_method"] == "none"
    write_csv(rows, str(tmp_path / "out" / "report.csv"))
    assert (tmp_path / "out" / "report.csv").read_text().splitlines()[0].startswith("task_idx,pcf_idx")
    # the README case at the golden depth and K: the reference's delta_exact
    g = golden["verify_pcf"][0]
    rows, _ = verify_pcf_records([{"a": "2", "b": "n**2", "limit": "2/(4 - pi)"}], depth=2000, K=32)
    want = float.fromhex(g["out"]["delta_exact"]) if rows[0]["delta_method"] == "crt" else \
        float.fromhex(g["out"]["delta_float"])
    assert rows[0]["delta_cuda_rns"] == f"{want:.5f}"
    assert rows[0]["p_bits"] == g["out"]["p_bits"]
