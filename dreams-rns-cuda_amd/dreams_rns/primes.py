"""RNS prime table (python/dreams_rns/runner.py:33-52).

The reference scans odd candidates downwards from 2^31 - 1, never below 2^30,
testing each by trial division, and returns the first K primes (fewer if the
range is exhausted). The same list is produced here with a deterministic
Miller-Rabin test (bases 2, 3, 5, 7 are exact below 3.2e9), which makes
K = 1024 instantaneous instead of minutes.
"""

from __future__ import annotations

from functools import lru_cache

import numpy as np

PRIME_MAX = (1 << 31) - 1
PRIME_MIN = 1 << 30


def is_prime(n: int) -> bool:
    if n < 2:
        return False
    for q in (2, 3, 5, 7, 11, 13):
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


@lru_cache(maxsize=None)
def _primes_tuple(K: int) -> tuple:
    out = []
    c = PRIME_MAX
    while len(out) < K and c >= PRIME_MIN:
        if is_prime(c):
            out.append(c)
        c -= 2
    return tuple(out)


def generate_rns_primes(K: int) -> np.ndarray:
    """First K primes below 2^31 (descending, >= 2^30) as int64."""
    return np.array(_primes_tuple(int(K)), dtype=np.int64)
