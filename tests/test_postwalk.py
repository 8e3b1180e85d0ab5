"""The step after the walk (no GPU): constant bank, delta helpers, the CRT
checks, and the vectorised sweep hit selection against the walk-by-walk
restatement of the reference loop (oracle/postwalk.py)."""

import math
import random
from fractions import Fraction

import mpmath as mp
import numpy as np
import pytest

from dreams_rns import _native as N
from dreams_rns.constants import (CONSTANTS_REGISTRY, compute_delta_against_constant, compute_match_digits,
                                  load_constants, match_against_constants)
from dreams_rns.engine import BatchResult
from dreams_rns.postwalk import (compute_delta, compute_delta_float, crt_overflowed, crt_ratio_ok, parse_limit,
                                 sweep_hits)
from dreams_rns.sweep import SweepBatch
from oracle import postwalk as opw
from oracle import walk as ow


@pytest.fixture(scope="module")
def bank():
    return load_constants(dps=60)


def test_bank_values(bank):
    assert [c["name"] for c in bank] == [c["name"] for c in CONSTANTS_REGISTRY]
    assert len(bank) == 40
    v = {c["name"]: c["value"] for c in bank}
    mp.mp.dps = 60
    assert abs(v["zeta2"] - mp.pi ** 2 / 6) < mp.mpf(10) ** -55
    assert abs(v["zeta4"] - mp.pi ** 4 / 90) < mp.mpf(10) ** -55
    assert abs(v["pi_sq6"] - v["zeta2"]) < mp.mpf(10) ** -55
    assert abs(v["phi"] - (1 + mp.sqrt(5)) / 2) < mp.mpf(10) ** -55
    assert abs(v["phi"] ** 2 - v["phi"] - 1) < mp.mpf(10) ** -55
    assert abs(v["4/pi"] * v["1/pi"] * v["pi2"] - 4) < mp.mpf(10) ** -55
    assert mp.nstr(v["catalan"], 30) == "0.915965594177219015054603514932"
    assert mp.nstr(v["euler_gamma"], 30) == "0.577215664901532860606512090082"
    assert mp.nstr(v["zeta3"], 30) == "1.20205690315959428539973816151"
    assert abs(mp.exp(v["ln2"]) - 2) < mp.mpf(10) ** -55 and abs(mp.log(v["e"]) - 1) < mp.mpf(10) ** -55
    # zeta(n) -> 1 from above, strictly decreasing
    z = [v[f"zeta{k}"] for k in range(2, 31)]
    assert all(a > b > 1 for a, b in zip(z, z[1:]))
    for c in bank:
        assert c["value_float"] == float(c["value"])


def test_match_helpers(bank):
    pi = float(mp.pi)
    name, d = compute_match_digits(pi + 1e-8, bank)
    assert name == "pi" and abs(d - (-math.log10(1e-8 / pi))) < 1e-6
    # near 1: tails decide, so zeta(12) is not taken for zeta(11)
    z11 = next(c["value_float"] for c in bank if c["name"] == "zeta11")
    name, d = compute_match_digits(1.0 + (z11 - 1.0) * (1 + 1e-6), bank)
    assert name == "zeta11" and 5.5 < d < 6.5
    assert compute_match_digits(pi, bank) == ("pi", 16.0)
    assert compute_match_digits(-1e6, [])[0] == "none"
    m = match_against_constants(1.6445, bank, 1e-2)
    assert [x["name"] for x in m] == ["zeta2", "pi_sq6"]  # equal values, bank order; phi is 0.026 away
    assert all(a["distance"] <= b["distance"] for a, b in zip(m, m[1:]))
    assert match_against_constants(100.0, bank) == []


def test_delta_closed_forms(bank):
    pi = next(c["value"] for c in bank if c["name"] == "pi")
    # 22/7 and 355/113 against pi: -(1 + ln|p/q - pi| / ln q), the error
    # taken exactly against a 40-digit decimal pi (Fractions), the logs in FP64
    pi40 = Fraction("3.141592653589793238462643383279502884197")
    for p, q in ((22, 7), (355, 113)):
        want = -(1 + math.log(float(abs(Fraction(p, q) - pi40))) / math.log(q))
        assert abs(compute_delta_against_constant(p, q, pi, 60) - want) < 1e-12
    d2 = compute_delta_against_constant(355, 113, pi, 60)
    assert abs(d2 - 2.2019587426065614) < 1e-12
    assert compute_delta_against_constant(-355, -113, pi, 60) == d2  # signs cancel; |q| enters the log
    assert compute_delta_against_constant(1, 0, pi, 60) == float("-inf")
    assert compute_delta_against_constant(3, 1, pi, 60) == float("-inf")  # ln 1 = 0
    tgt = parse_limit("2/(4 - pi)", 60)
    assert abs(tgt - 2 / (4 - mp.pi)) < mp.mpf(10) ** -55
    assert parse_limit("sqrt(-1)") is None and parse_limit("foo(") is None
    assert compute_delta(355, 113, pi, 60) == d2
    assert compute_delta(1, 0, pi) == float("-inf")
    assert compute_delta_float(3.0, 3.0, 1.0, 1.0) == float("inf")
    assert compute_delta_float(3.1, float(pi), 0.0, 1.0) == float("-inf")  # log q = 0
    assert abs(compute_delta_float(3.14, float(pi), 2.0, 1.0) - (-(1 + math.log(abs(3.14 - float(pi))) / 2.0))) < 1e-15


def test_crt_checks():
    big = 1 << 2000
    # mpmath handles integers beyond FP64; the sweep's float() check does not
    assert crt_overflowed(3 * big, 2 * big, 1.5) is False
    assert crt_ratio_ok(3 * big, 2 * big, 1.5) is False
    assert crt_ratio_ok(3, 2, 1.5) is True and crt_overflowed(3, 2, 1.5) is False
    assert crt_ratio_ok(3, 2, 1.6) is False and crt_overflowed(3, 2, 1.6) is True
    assert crt_ratio_ok(1, 0, 1.0) is False and crt_overflowed(1, 0, 1.0) is True
    assert crt_ratio_ok(1, 10 ** 12, 0.0) is True  # absolute test near 0
    assert crt_overflowed(1, 2, float("nan")) is True
    # the largest double is 2^1024 - 2^971; 2^1024 - 2^970 (the halfway point) rounds
    # to 2^1024, so the conversion itself overflows
    assert crt_ratio_ok((1 << 1024) - (1 << 971), 1 << 1023, 2.0) is True
    assert crt_ratio_ok((1 << 1024) - (1 << 970), 1 << 1023, 2.0) is False


def _row(x: int, primes):
    res = [x % int(p) for p in primes]
    mag = abs(x)
    limbs = [(mag >> (32 * t)) & 0xFFFFFFFF for t in range(len(primes))]
    return res, limbs, 1 if x < 0 else 0


def _fake_batch(rows, K):
    """rows: (p_big, q_big, p_float, q_float, log_scale, shift indices)."""
    primes = ow.generate_rns_primes(K)
    n = len(rows)
    resid = np.zeros((n, 2, K), dtype=np.int64)
    limbs = np.zeros((n, 2, K), dtype=np.uint32)
    neg = np.zeros((n, 2), dtype=np.uint8)
    out = np.zeros(n, dtype=N.OUT_DTYPE)
    for i, (pb, qb, pf, qf, ls, _) in enumerate(rows):
        for v, x in enumerate((pb, qb)):
            r, lm, s = _row(x, primes)
            resid[i, v] = r
            limbs[i, v] = lm
            neg[i, v] = s
        out[i]["p_float"], out[i]["q_float"], out[i]["log_scale"] = pf, qf, ls
    res = BatchResult(residues=resid, out=out, primes=primes, Ks=np.full(n, K, dtype=np.int32), limbs=limbs, neg=neg)
    return SweepBatch(spec_idx=np.array([i % 2 for i in range(n)], dtype=np.int32),
                      traj_idx=np.array([i // 2 for i in range(n)], dtype=np.int32),
                      offsets=[[1 + i, 2] for i in range(n)], shift_rows=[r[5] for r in rows], result=res, crt=True)


def _oracle_hits(batch, specs, trajs, bank, proximity, dps):
    out = []
    for i in range(len(batch)):
        d = batch.result.walk_dict(i)
        h = opw.walk_record(d, bank, proximity, dps)
        if h is None:
            continue
        si, ti = int(batch.spec_idx[i]), int(batch.traj_idx[i])
        for s in batch.shift_rows[i]:
            out.append((si, ti, s, {"cmf_name": specs[si]["name"], "spec_hash": specs[si]["spec_hash"],
                                    "rank": specs[si]["rank"], "traj_idx": ti, "traj": list(trajs[ti]),
                                    "shift_idx": s, "shift_offsets": list(batch.offsets[i]), **h}))
    out.sort(key=lambda t: t[:3])
    return [t[3] for t in out]


def test_sweep_hits_match_the_reference_loop(bank):
    rng = random.Random(3)
    cv = {c["name"]: c["value_float"] for c in bank}
    K = 32
    M = math.prod(int(p) for p in ow.generate_rns_primes(K))

    def rnd(bits):
        return rng.getrandbits(bits) * (1 if rng.random() < 0.5 else -1)

    k = 10 ** 18 + 7
    rows = [
        (355 * k, 113 * k, 355.0, 113.0, 0.0, [0, 5]),                 # exact CRT, proximity to pi
        (14 * 10 ** 20 + 3, 10 ** 21, 1.4, 1.0, 0.0, [1]),             # exact CRT, no hit (ln q large)
        (7, 5, 1.4, 1.0, 0.0, [2]),                                    # exact CRT, delta > 0 against sqrt2
        (rnd(900), rnd(900), 1.3, 1.0, 40.0, [3]),                     # overflowed, no hit
        (rnd(980), rnd(980), cv["zeta3"] + 1e-5, 1.0, 40.0, [4, 9]),   # overflowed, proximity
        (rnd(900), rnd(900), cv["zeta5"], 1.0, 40.0, [6]),             # proximity at distance 0: delta -inf
        (rnd(900), rnd(900), cv["ln2"] + 0.002, 1.0, 2.0, [7]),        # float delta > 0, no proximity
        (rnd(900), rnd(900), cv["ln2"] + 0.002, 1.0, 60.0, [8]),       # float delta < 0
        (rnd(900), 0, 1.0, 0.0, 1.0, [10]),                            # estimate not finite: skipped
        (rnd(900), 0, 2.5, 1.0, 1.0, [11]),                            # q = 0: float path
        (-(355 * k), -(113 * k), -355.0, -113.0, 0.0, [12]),            # negative integers
        (rnd(900), rnd(900), 2 * (cv["e"] + 0.0009), 2.0, 3.0, [13]),  # proximity, q_float 2
        (M // 2 - 5, M // 2 - 7, 1.0, 1.0, 0.0, [14]),                 # large exact ratio, near 1
    ]
    batch = _fake_batch(rows, K)
    specs = [{"name": "A", "spec_hash": "h0", "rank": 4}, {"name": "B", "spec_hash": "h1", "rank": 3}]
    trajs = [[1, 0], [0, 1], [1, 1], [2, 1], [1, 2], [3, 1], [1, 3]]
    got = sweep_hits(batch, specs, trajs, bank, proximity=1e-3, dps=60)
    ref = _oracle_hits(batch, specs, trajs, bank, 1e-3, 60)
    assert [h["shift_idx"] for h in got] == [h["shift_idx"] for h in ref]
    for g, r in zip(got, ref):
        for key, val in r.items():
            assert g[key] == val or (isinstance(val, float) and math.isnan(val) and math.isnan(g[key])), (key, g, r)
    hit_shifts = {h["shift_idx"] for h in got}
    assert {0, 5, 2, 4, 9, 6, 7, 12, 13} <= hit_shifts and not ({1, 3, 8, 10} & hit_shifts)
    # the screen changes nothing: every row through the reference computation
    assert sweep_hits(batch, specs, trajs, bank, proximity=1e-3, dps=60, margin=1e9) == got


def test_sweep_hits_random_rows(bank):
    # many walks near and far from the bank, overflowed and exact
    rng = random.Random(11)
    vals = [c["value_float"] for c in bank]
    rows = []
    for i in range(300):
        c = rng.choice(vals)
        est = c + rng.choice([0.0, 1e-9, 1e-4, 2e-3, 0.05, 0.7]) * rng.choice([-1, 1])
        if rng.random() < 0.3:
            q = rng.randint(2, 10 ** 6)
            p = round(est * q)
            rows.append((p, q, float(p), float(q), 0.0, [i]))
        else:
            rows.append((rng.getrandbits(950), rng.getrandbits(950) | 1, est, 1.0, rng.uniform(0.5, 80), [i]))
    batch = _fake_batch(rows, 32)
    specs = [{"name": "A", "spec_hash": "h0", "rank": 4}, {"name": "B", "spec_hash": "h1", "rank": 3}]
    trajs = [[i, 1] for i in range(200)]
    got = sweep_hits(batch, specs, trajs, bank, proximity=1e-3, dps=60)
    ref = _oracle_hits(batch, specs, trajs, bank, 1e-3, 60)
    assert len(got) == len(ref) and len(ref) > 50
    for g, r in zip(got, ref):
        assert {k: g[k] for k in r} == r
