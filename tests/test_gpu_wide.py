"""Walks and CRT over the reference's whole prime range: K above 1024 primes
(euler2ai_full_verify.py's automatic K reaches 2000-4300 at depth 2000,
euler2ai_full_verify.py:40-92, 331-333), the wide Garner kernel against the
textbook CRT and round trips, crt_reconstruct on any odd primes with the
reference's argument rules (runner.py:57-76), and the Euler2AI report at
that script's defaults without ERROR rows."""

import math

import numpy as np
import pytest

import dreams_rns as D
from dreams_rns.engine import BatchPlan, pcf_spec
from oracle import cref
from oracle import walk as ow

pytestmark = pytest.mark.gpu


def _crt_rows(ctx, res, K, stride=None):
    n = res.shape[0]
    stride = stride or K
    limbs = np.zeros((n, stride), dtype=np.uint32)
    neg = np.zeros(n, dtype=np.uint8)
    r = np.zeros((n, stride), dtype=np.int64)
    r[:, :K] = res
    ctx._check(ctx.lib.drx_crt(ctx.h, r.ctypes.data, n, K, stride, limbs.ctypes.data, neg.ctypes.data))
    out = []
    for i in range(n):
        mag = int.from_bytes(limbs[i].tobytes(), "little")
        assert not limbs[i, K:].any()
        out.append(-mag if neg[i] else mag)
    return out


@pytest.mark.parametrize("K", [257, 300, 1000, 1024, 1025, 2001, 4096, 8192])
def test_wide_garner_rows_equal_textbook(gpu_ctx, K):
    # fewer than 64 rows: the wide Garner kernel (one workgroup per row)
    gpu_ctx.ensure_primes(K)
    pr = [int(p) for p in gpu_ctx.primes[:K]]
    M = math.prod(pr)
    H = (M - 1) // 2
    rng = np.random.default_rng(K)
    vals = [0, 1, -1, H, -H, H - 1, -(H - 1), M // 3, -(M // 7), 2 ** 31, -(2 ** 200)]
    vals += [int.from_bytes(rng.bytes(8 * K), "little") % M - H for _ in range(6)]
    rows = [[v % p for p in pr] for v in vals]
    print("wide crt K", K, "rows", len(rows), flush=True)
    got = _crt_rows(gpu_ctx, np.array(rows, dtype=np.int64), K, stride=K + 3)
    assert got == vals
    # random residue rows against the reference's iterative CRT
    rows = [[int(rng.integers(0, p)) for p in pr] for _ in range(3)]
    got = _crt_rows(gpu_ctx, np.array(rows, dtype=np.int64), K)
    for g, row in zip(got, rows):
        x, MM = ow.crt_reconstruct(row, pr)
        assert MM == M and g == ow.centered(x, M)


def test_matrix_and_wide_paths_agree(gpu_ctx):
    # 80 rows of 600 primes take the matrix-core CRT, 8 rows the wide kernel
    K = 600
    gpu_ctx.ensure_primes(K)
    pr = [int(p) for p in gpu_ctx.primes[:K]]
    rng = np.random.default_rng(3)
    res = np.array([[int(rng.integers(0, p)) for p in pr] for _ in range(80)], dtype=np.int64)
    many = _crt_rows(gpu_ctx, res, K)
    few = _crt_rows(gpu_ctx, res[:8], K)
    assert many[:8] == few


def test_crt_reconstruct_takes_any_odd_primes(gpu_ctx):
    # the reference's own call shapes (runner.py:57-76)
    assert D.crt_reconstruct([2, 3], [5, 7]) == (17, 35)
    assert D.crt_reconstruct([7], [5]) == (7, 5)  # the first residue is taken as given
    assert D.crt_reconstruct([12, 3], [5, 7]) == (17, 35)
    assert D.crt_reconstruct([2, 3], [5, 7, 11]) == (17, 35)  # extra primes are not used
    x, M = D.crt_reconstruct([1, 2, 3, 4], [3, 5, 7, 11])
    assert M == 1155 and [x % p for p in (3, 5, 7, 11)] == [1, 2, 3, 4] and 0 <= x < M
    big = [2147483647, 65537, 101, 3]
    for v in (0, 1, 12345678, math.prod(big) - 1):
        assert D.crt_reconstruct([v % p for p in big], big) == (v, math.prod(big))
    with pytest.raises(IndexError):
        D.crt_reconstruct([1, 2, 3], [5, 7])
    with pytest.raises(ValueError):
        D.crt_reconstruct([1, 2], [9, 7])  # not a prime
    with pytest.raises(ValueError):
        D.crt_reconstruct([1, 1], [7, 7])  # not coprime


def test_pcf_walk_at_2001_primes_equals_c_oracle(gpu_ctx):
    # the README PCF at depth 2000 with the K its estimate gives (2001)
    from dreams_rns.euler2ai import estimate_K

    K = estimate_K("2", "n**2", 2000)
    assert K == 2001
    prog = D.compile_pcf("2", "n**2")
    got = D.run_pcf_walk(prog, 2, 2000, K)
    primes = D.generate_rns_primes(K)
    rr, pf, ls = cref.pcf_walk(prog, 2, 2000, primes)
    assert got["p_residues"].tolist() == rr[0, 1].tolist()
    assert got["q_residues"].tolist() == rr[1, 1].tolist()
    assert got["p_float"] == pf[0, 1] and got["q_float"] == pf[1, 1]
    assert abs(got["log_scale"] - ls) <= 1e-12 * abs(ls)
    # CRT integers of the same walk through the plan, against Python CRT
    with BatchPlan([pcf_spec(prog, 2, 2000, K, 1)], crt=True) as plan:
        p_big, q_big = plan.run().crt_values(0)
    pl = [int(p) for p in primes]
    xp, M = ow.crt_reconstruct(rr[0, 1].tolist(), pl)
    xq, _ = ow.crt_reconstruct(rr[1, 1].tolist(), pl)
    assert (p_big, q_big) == (ow.centered(xp, M), ow.centered(xq, M))
    assert D.crt_reconstruct(got["q_residues"], got["primes"]) == (xq, M)


def test_widest_walk_and_one_more_prime(gpu_ctx):
    prog = D.compile_pcf("3*n + 1", "-2*n**2")
    K = 8192
    got = D.run_pcf_walk(prog, 1, 60, K)
    rr, _, _ = cref.pcf_walk(prog, 1, 60, D.generate_rns_primes(K))
    assert got["q_residues"].tolist() == rr[1, 1].tolist()
    with pytest.raises(ValueError, match="8193"):
        BatchPlan([pcf_spec(prog, 1, 10, 8193, 1)])


def test_euler2ai_defaults_give_no_error_rows(gpu_ctx):
    # depth 2000 with the automatic K (euler2ai_full_verify.py:192-193, 331-333)
    from dreams_rns.euler2ai import estimate_K, verify_pcf_records
    from dreams_rns.postwalk import compute_delta, compute_delta_float, crt_overflowed, parse_limit

    recs = [{"a": "2", "b": "n**2", "limit": "2/(4 - pi)", "delta": -1.0},
            {"a": "3*n + 1", "b": "-2*n**2", "limit": "1/log(2)", "delta": 0.2},
            {"a": "5*n**2 + 3", "b": "-n**4", "limit": "", "delta": 1.0},
            {"a": "n + 1", "b": "n", "limit": "E - 1", "delta": 0.0},
            {"a": "7*n**3 - 2*n", "b": "-12*n**6 + 5", "limit": "1", "delta": -0.5}]
    depth = 2000
    rows, summ = verify_pcf_records(recs, depth=depth, K=0, dps=200)
    print("euler2ai defaults", summ, [r["K"] for r in rows], flush=True)
    assert summ["errors"] == 0 and all(r["match"] not in ("ERROR", "SKIP") for r in rows)
    assert [r["K"] for r in rows] == [2001, 2041, 3942, 926, 5893]  # four of them above 1024 primes
    for r, rec in zip(rows, recs):
        K = estimate_K(rec["a"], rec["b"], depth)
        assert r["K"] == K
        prog = D.compile_pcf(rec["a"], rec["b"])
        primes = [int(p) for p in D.generate_rns_primes(K)]
        rr, pf, ls = cref.pcf_walk(prog, D.pcf_initial_values(rec["a"]), depth, primes)
        xp, M = ow.crt_reconstruct(rr[0, 1].tolist(), primes)
        xq, _ = ow.crt_reconstruct(rr[1, 1].tolist(), primes)
        p_big, q_big = ow.centered(xp, M), ow.centered(xq, M)
        est = pf[0, 1] / pf[1, 1] if abs(pf[1, 1]) > 1e-300 else float("nan")
        assert r["est_float"] == (f"{est:.10f}" if math.isfinite(est) else str(est))
        assert r["p_bits"] == M.bit_length()
        if not rec["limit"]:
            assert r["delta_method"] == "none"
            continue
        target = parse_limit(rec["limit"], 200)
        if crt_overflowed(p_big, q_big, est):
            method, delta = "float", compute_delta_float(est, float(target), ls, pf[1, 1])
        else:
            method, delta = "crt", compute_delta(p_big, q_big, target, 200)
        assert r["delta_method"] == method
        assert r["delta_cuda_rns"] == (f"{delta:.5f}" if math.isfinite(delta) else str(delta))
