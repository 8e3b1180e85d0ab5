"""The matrix-core CRT pipeline (tests/crtmodel.py) against exact CRT: the
integer decomposition the device kernels use is pinned here on the CPU."""

import random

import numpy as np

import crtmodel as cm
from oracle import walk as ow


def _centered_crt(res_row, primes):
    x, M = ow.crt_reconstruct([int(r) for r in res_row], [int(p) for p in primes])
    return ow.centered(x, M)


def test_balanced_digits_round_trip():
    rng = random.Random(1)
    for _ in range(200):
        v = rng.randrange(-(1 << 30), 1 << 30)
        d = cm.balanced_digits(v, 4)
        assert all(-128 <= x <= 127 for x in d)
        assert sum(x << (8 * t) for t, x in enumerate(d)) == v


def test_model_equals_exact_crt():
    rng = random.Random(7)
    for K in (3, 17, 64, 100):
        primes = ow.generate_rns_primes(K)
        M = 1
        for p in primes:
            M *= int(p)
        vals = [0, 1, -1, (M - 1) // 2, -((M - 1) // 2), (M - 1) // 2 - 1, M // 3, -(M // 5)]
        vals += [rng.randrange(-(M // 2), M // 2) for _ in range(40)]
        res = np.array([[v % int(p) for p in primes] for v in vals], dtype=np.int64)
        got = cm.crt_rows(res, primes)
        assert got == vals, K
        # random residue rows (any x in [0, M))
        res = np.array([[rng.randrange(int(p)) for p in primes] for _ in range(30)], dtype=np.int64)
        assert cm.crt_rows(res, primes) == [_centered_crt(r, primes) for r in res]


def test_wide_garner_model_equals_exact_crt():
    # the wide Garner kernel's steps (tests/crtmodel.py wide_garner_row) on
    # layouts with one and several digits per thread, against exact CRT;
    # small odd primes included (the CRT accepts any odd primes below 2^31)
    rng = random.Random(11)
    cases = [(ow.generate_rns_primes(K), nt, r) for K, nt, r in ((1, 4, 1), (5, 4, 2), (13, 4, 4), (31, 8, 4), (40, 16, 4))]
    cases.append(([3, 5, 7, 11, 13, 101, 65537], 2, 4))
    for primes, nt, r in cases:
        pr = [int(p) for p in primes]
        M = 1
        for p in pr:
            M *= p
        H = (M - 1) // 2
        vals = [0, 1, -1, H, -H, H - 1, -(H - 1), M // 3, -(M // 7)] + [rng.randrange(-H, H + 1) for _ in range(20)]
        for v in vals:
            mag, neg = cm.wide_garner_row([v % p for p in pr], pr, nt, r)
            assert (-mag if neg else mag) == v, (pr[:3], nt, r, v)
        for _ in range(10):
            row = [rng.randrange(p) for p in pr]
            mag, neg = cm.wide_garner_row(row, pr, nt, r)
            assert (-mag if neg else mag) == _centered_crt(row, pr)
