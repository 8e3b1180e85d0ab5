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
M_i in balanced base-256 digits (columns t), int8 x int8 -> int32, and
S = sum_{b,t} C[b][t] 256^(b+t).
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
    C = np.einsum("bji,ti->btj", A, bt.astype(np.int64))  # [b][t][j], each |.| < 2^26
    assert np.abs(C).max() < 2 ** 26
    H = (M - 1) // 2
    out = []
    for j in range(n):
        S = 0
        for b in range(4):
            for t in range(C.shape[1]):
                S += int(C[b, t, j]) << (8 * (b + t))
        x = S - kq[j] * M
        if x > H:
            x -= M
        elif x < -H:
            x += M
        assert -H <= x <= H
        out.append(x)
    return out
