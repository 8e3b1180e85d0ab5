"""CPU model of the matrix-core CRT (test infrastructure): the same integer
pipeline as the device kernels k_crt_digits / k_crt_gemm / k_crt_fold, in
numpy and Python integers, so every step can be checked against exact CRT.

For residues r_i of x modulo the first K primes (M = prod p_i, M_i = M / p_i,
w_i = M_i^-1 mod p_i):
  c_i = r_i w_i mod p_i, centred into (-p_i/2, p_i/2);
  S   = sum_i c_i M_i  (|S| < K M / 2), S = x mod M;
  k   = rint(sum_i c_i / p_i) in FP64, so S - k M is the centred x up to one
        multiple of M, fixed by an exact comparison with (M - 1) / 2.
S is a matrix product: c_i in 4 balanced base-256 digits (planes b) times
M_i in balanced base-256 digits (columns t), int8 x int8 -> int32. The
planes are one reduction, D[t] = sum_b sum_i A[b][i] B[t - b][i], the
coefficient of 256^t in S = sum_t D[t] 256^t, and the fold builds 32-bit
limb l from D[4l .. 4l + 3].
"""

from __future__ import annotations

from typing import List, Tuple

import numpy as np


def balanced_digits(v: int, n: int) -> List[int]:
    """n digits in [-128, 127] with v = sum d_t 256^t (v must fit)."""
    out = []
    for _ in range(n):
        d = v & 255
        if d >= 128:
            d -= 256
        out.append(d)
        v = (v - d) >> 8
    assert v == 0, "value does not fit the digits"
    return out


def tables(primes) -> Tuple[np.ndarray, List[int], int, int]:
    """B^T [T][K] int8 (balanced base-256 digits of M_i), w_i, M, limbs of M."""
    pr = [int(p) for p in primes]
    M = 1
    for p in pr:
        M *= p
    L = (M.bit_length() + 31) // 32
    T = 4 * L + 1
    bt = np.zeros((T, len(pr)), dtype=np.int8)
    w = []
    for i, p in enumerate(pr):
        Mi = M // p
        w.append(pow(Mi % p, p - 2, p))
        bt[:, i] = balanced_digits(Mi, T)
    return bt, w, M, L


def crt_rows(res: np.ndarray, primes) -> List[int]:
    """Centred CRT of every row of res [n][K] through the model pipeline."""
    pr = [int(p) for p in primes]
    K = len(pr)
    bt, w, M, L = tables(pr)
    n = res.shape[0]
    A = np.zeros((4, n, K), dtype=np.int64)
    kq = []
    for j in range(n):
        f = 0.0
        for i, p in enumerate(pr):
            c = int(res[j, i]) * w[i] % p
            if c > (p - 1) // 2:
                c -= p
            A[:, j, i] = balanced_digits(c, 4)
            f += c * (1.0 / p)
        kq.append(int(np.rint(f)))
    T = bt.shape[0]
    C = np.einsum("bji,ti->btj", A, bt.astype(np.int64))  # [b][t][j]
    D = np.zeros((T + 3, n), dtype=np.int64)  # D[t] = sum_b C[b][t - b]
    for b in range(4):
        D[b:b + T] += C[b]
    assert np.abs(D).max() < 2 ** 26
    NL = L + 2
    Dp = np.zeros((4 * NL, n), dtype=np.int64)
    Dp[:min(4 * NL, T + 3)] = D[:4 * NL]
    H = (M - 1) // 2
    out = []
    for j in range(n):
        S = 0
        for l in range(NL):  # limb l from D[4l .. 4l + 3]
            v = sum(int(Dp[4 * l + q, j]) << (8 * q) for q in range(4))
            assert abs(v) < 2 ** 53
            S += v << (32 * l)
        x = S - kq[j] * M
        if x > H:
            x -= M
        elif x < -H:
            x += M
        assert -H <= x <= H
        out.append(x)
    return out


def wide_garner_w(primes) -> List[int]:
    """w_i = (p_0 ... p_{i-1})^-1 mod p_i, the table of the wide Garner kernel
    (w_0 = 1)."""
    pr = [int(p) for p in primes]
    out = []
    for i, p in enumerate(pr):
        prod = 1
        for q in pr[:i]:
            prod = prod * q % p
        out.append(pow(prod, p - 2, p))
    return out


def wide_garner_row(res_row, primes, NT: int, R: int) -> Tuple[int, int]:
    """(magnitude, negative) of the centred CRT of one row, computed the way
    k_garner_wide does: digit i in thread i % NT, register i / NT; the
    u/Q recurrence for the mixed-radix digits; sign and negation from the
    digits; Horner on redundant limbs (thread t holds limbs t R .. t R + R - 1,
    each kept below 2^33, one neighbour word per step); final carry pass."""
    pr = [int(p) for p in primes]
    K = len(pr)
    assert K <= NT * R
    w = wide_garner_w(pr)
    u = [int(res_row[i]) for i in range(K)]
    Q = [1] * K
    v = [0] * K
    v[0] = u[0] % pr[0]
    for j in range(K - 1):
        for i in range(j + 1, K):  # every later digit, in whichever thread holds it
            u[i] = (u[i] - v[j] * Q[i]) % pr[i]
            Q[i] = Q[i] * pr[j] % pr[i]
        v[j + 1] = u[j + 1] * w[j + 1] % pr[j + 1]
    h = [(p - 1) // 2 for p in pr]
    diff = [i for i in range(K) if v[i] != h[i]]
    neg = bool(diff) and v[diff[-1]] > h[diff[-1]]
    if neg:
        z = min(i for i in range(K) if v[i] != 0)
        v = [0 if i < z else (pr[i] - v[i] if i == z else pr[i] - 1 - v[i]) for i in range(K)]
    NL = NT * R
    L = [0] * NL
    L[0] = v[K - 1]
    for i in range(K - 2, -1, -1):
        z = [L[t] * pr[i] for t in range(NL)]
        assert max(z) < 1 << 64
        lo = [x & 0xFFFFFFFF for x in z]
        hi = [x >> 32 for x in z]
        assert hi[NL - 1] == 0  # the dropped top word is always zero
        new = []
        for t in range(NL):
            tid, b = divmod(t, R)
            if b == 0:  # the neighbouring thread's top word (thread 0: the digit)
                new.append(lo[t] + (hi[t - 1] if tid else v[i]))
            else:
                new.append(lo[t] + hi[t - 1])
        L = new
        assert max(L) < 1 << 33
    carry, mag = 0, 0
    for t in range(K):
        s = L[t] + carry
        mag |= (s & 0xFFFFFFFF) << (32 * t)
        carry = s >> 32
    assert carry == 0 and all(x == 0 for x in L[K:])
    return mag, neg
