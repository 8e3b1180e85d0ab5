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
