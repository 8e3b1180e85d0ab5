/* cwalk.c -- C restatement of the reference walk, TEST INFRASTRUCTURE ONLY.
 *
 * Follows python/dreams_rns/runner.py:198-260 (bytecode over the residues),
 * runner.py:263-316 / cmf_walk.py:261-313 (FP64 bytecode), cmf_walk.py:174-188
 * (matrix product reduced after every term) and runner.py:393-398 (shadow:
 * P_f @ M_f, then divide by max|P_f| when it exceeds 1e10 and add its log).
 * The FP64 product is written as the sequential FMA chain
 * acc = a0*b0; acc = fma(ak, bk, acc) that OpenBLAS dgemm evaluates on the
 * machine that produced the golden vectors; POW2 / POW3 call pow() like
 * Python's float ** int. Only the checker in tests/, smoke() and bench.py's
 * CPU baseline use this file.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { NOP, LOAD_X, LOAD_C, LOAD_N, ADD, SUB, MUL, NEG, POW2, POW3, INV, STORE, END };
#define NREG 512

static uint64_t powmod(uint64_t a, uint64_t e, uint64_t p) {
  uint64_t r = 1 % p;
  a %= p;
  while (e) {
    if (e & 1) r = r * a % p;
    a = a * a % p;
    e >>= 1;
  }
  return r;
}

static uint64_t mod128(__int128 v, uint64_t p) {
  __int128 r = v % (__int128)p;
  return (uint64_t)(r < 0 ? r + (__int128)p : r);
}

/* volatile exponents keep the compiler from turning pow into multiplies */
static volatile double two = 2.0, three = 3.0;

/* Walk one program. consts are int64 values; axes holds (sn, sd, tn, td)
 * per axis; init_res [m*m][K] canonical; init_f [m*m]. Outputs the final
 * residues [m*m][K], the final P_f [m*m] and log_scale. Returns 0 or -1. */
int cw_walk(int m, int dim, const uint32_t* ops, int n_ops, const int64_t* consts, int n_const,
            const int64_t* primes, int K, int64_t depth, const int64_t* init_res, const double* init_f,
            const int64_t* axes, int64_t* out_res, double* out_f, double* out_log_scale) {
  const int mm = m * m;
  uint64_t* P = malloc(sizeof(uint64_t) * mm * K);
  uint64_t* Mr = malloc(sizeof(uint64_t) * mm);
  uint64_t* C = malloc(sizeof(uint64_t) * mm);
  uint64_t* R = malloc(sizeof(uint64_t) * NREG);
  double* F = malloc(sizeof(double) * NREG);
  double Pf[64], Mf[64], Cf[64];
  if (!P || !Mr || !C || !R || !F || mm > 64) return -1;
  for (int i = 0; i < mm * K; ++i) P[i] = (uint64_t)init_res[i];
  for (int i = 0; i < mm; ++i) Pf[i] = init_f[i];
  double log_scale = 0.0;
  for (int64_t s = 0; s < depth; ++s) {
    /* residues: prime by prime */
    for (int k = 0; k < K; ++k) {
      const uint64_t p = (uint64_t)primes[k];
      memset(R, 0, sizeof(uint64_t) * NREG);
      memset(Mr, 0, sizeof(uint64_t) * mm);
      for (int i = 0; i < n_ops; ++i) {
        const uint32_t op = ops[4 * i], a0 = ops[4 * i + 1], a1 = ops[4 * i + 2], a2 = ops[4 * i + 3];
        if (op == END) break;
        switch (op) {
          case LOAD_X: {
            const int64_t* q = axes + 4 * a0;
            const __int128 num = (__int128)q[0] * q[3] + (__int128)s * q[2] * q[1];
            const __int128 den = (__int128)q[1] * q[3];
            const uint64_t inv = den == 1 ? 1 : powmod(mod128(den, p), p - 2, p);
            R[a1] = mod128(num, p) * inv % p;
            break;
          }
          case LOAD_C: R[a1] = mod128(consts[a0], p); break;
          case LOAD_N: R[a0] = (uint64_t)s % p; break;
          case ADD: R[a2] = (R[a0] + R[a1]) % p; break;
          case SUB: R[a2] = (R[a0] + p - R[a1]) % p; break;
          case MUL: R[a2] = R[a0] * R[a1] % p; break;
          case NEG: R[a1] = (p - R[a0]) % p; break;
          case POW2: R[a1] = R[a0] * R[a0] % p; break;
          case POW3: R[a1] = R[a0] * R[a0] % p * R[a0] % p; break;
          case INV: R[a1] = powmod(R[a0], p - 2, p); break;
          case STORE: Mr[a1 * m + a2] = R[a0] % p; break;
          default: break;
        }
      }
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
          uint64_t acc = 0;
          for (int t = 0; t < m; ++t) acc = (acc + P[(i * m + t) * K + k] * Mr[t * m + j]) % p;
          C[i * m + j] = acc;
        }
      for (int e = 0; e < mm; ++e) P[e * K + k] = C[e];
    }
    /* FP64 shadow */
    for (int i = 0; i < NREG; ++i) F[i] = 0.0;
    for (int i = 0; i < mm; ++i) Mf[i] = 0.0;
    for (int i = 0; i < n_ops; ++i) {
      const uint32_t op = ops[4 * i], a0 = ops[4 * i + 1], a1 = ops[4 * i + 2], a2 = ops[4 * i + 3];
      if (op == END) break;
      switch (op) {
        case LOAD_X: {
          const int64_t* q = axes + 4 * a0;
          const int64_t num = q[0] * q[3] + s * q[2] * q[1];
          F[a1] = (double)num / (double)(q[1] * q[3]);
          break;
        }
        case LOAD_C: F[a1] = (double)consts[a0]; break;
        case LOAD_N: F[a0] = (double)s; break;
        case ADD: F[a2] = F[a0] + F[a1]; break;
        case SUB: F[a2] = F[a0] - F[a1]; break;
        case MUL: F[a2] = F[a0] * F[a1]; break;
        case NEG: F[a1] = -F[a0]; break;
        case POW2: F[a1] = pow(F[a0], two); break;
        case POW3: F[a1] = pow(F[a0], three); break;
        case INV: F[a1] = fabs(F[a0]) > 1e-300 ? 1.0 / F[a0] : 0.0; break;
        case STORE: Mf[a1 * m + a2] = F[a0]; break;
        default: break;
      }
    }
    double mx = 0.0;
    int nan = 0;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) {
        double acc = Pf[i * m] * Mf[j];
        for (int t = 1; t < m; ++t) acc = fma(Pf[i * m + t], Mf[t * m + j], acc);
        Cf[i * m + j] = acc;
        const double a = fabs(acc);
        if (a != a) nan = 1;
        if (a > mx) mx = a;
      }
    if (!nan && mx > 1e10) {
      for (int e = 0; e < mm; ++e) Cf[e] /= mx;
      log_scale += log(mx);
    }
    memcpy(Pf, Cf, sizeof(double) * mm);
  }
  for (int i = 0; i < mm * K; ++i) out_res[i] = (int64_t)P[i];
  for (int i = 0; i < mm; ++i) out_f[i] = Pf[i];
  *out_log_scale = log_scale;
  free(P);
  free(Mr);
  free(C);
  free(R);
  free(F);
  return 0;
}
