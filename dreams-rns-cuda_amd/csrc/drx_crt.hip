// Matrix-core CRT for wide prime sets (internal to libdreams_hip).
//
// For residues r_i of x modulo the first K primes (M = prod p_i,
// M_i = M / p_i, w_i = M_i^-1 mod p_i) the centred x is
//   S - k M,  S = sum_i c_i M_i,  c_i = r_i w_i mod p_i centred into
//   (-p_i/2, p_i/2),  k = rint(sum_i c_i / p_i),
// up to one multiple of M that an exact comparison with (M - 1) / 2 settles.
// S is an integer matrix product: the c_i of a row in four balanced base-256
// digits (planes b) times the M_i in balanced base-256 digits (columns t),
// int8 x int8 -> int32 on the MFMA units. The planes are one reduction:
// D[t] = sum_b sum_i A[b][i] B[t - b][i] is the coefficient of 256^t in
// S = sum_t D[t] 256^t (|D| < 4 K 2^14 <= 2^26), folded into 32-bit limbs by
// one lane per row. Work per row: 4 K T multiply-adds with T ~ 4 * 31 K / 32, against
// Garner's K^2 / 2 dependent modular steps; the matrix units make it the
// faster one for wide K.
#include "drx_crt.h"

#include "drx_device.h"

namespace drx {

typedef int v16i __attribute__((ext_vector_type(16)));
typedef int v4i __attribute__((ext_vector_type(4)));

// One block per (padded) row: the row's digit planes A[b][j][0..Kp) and
// k_j = rint(sum c_i / p_i). Rows past n get zero planes (the GEMM reads the
// whole padded tile).
__global__ void __launch_bounds__(256) k_crt_digits(const CrtArgs g) {
  __shared__ double red_buf[256];
  const int j = blockIdx.x, tid = threadIdx.x;
  const size_t plane = (size_t)g.npad * g.Kp;
  int8_t* A = g.A + (size_t)j * g.Kp;
  if (j >= g.n) {  // uniform over the block
    for (int i = tid; i < g.Kp; i += blockDim.x)
#pragma unroll
      for (int b = 0; b < 4; ++b) A[b * plane + i] = 0;
    return;
  }
  const int64_t* r = g.res + (size_t)g.rows[j] * g.res_stride;
  double f = 0.0;
  for (int i = tid; i < g.Kp; i += blockDim.x) {
    int c = 0;
    if (i < g.K) {
      const double p = g.prime_p[i], pinv = g.prime_pinv[i];
      double x = canon(mm((double)r[i], g.w[i], p, pinv), p, pinv);
      if (x > 0.5 * (p - 1.0)) x -= p;  // centred: |x| < 2^30
      c = (int)x;
      f = __builtin_fma(x, pinv, f);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int d = (int)(int8_t)(uint8_t)(c & 255);  // balanced digit in [-128, 127]
      A[b * plane + i] = (int8_t)d;
      c = (c - d) >> 8;  // exact: c - d is a multiple of 256
    }
  }
  red_buf[tid] = f;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) red_buf[tid] += red_buf[tid + s];
    __syncthreads();
  }
  if (tid == 0) {
    const double k = __builtin_rint(red_buf[0]);
    g.kq[j] = (int32_t)k;
    g.fd[j] = red_buf[0] - k;  // x / M up to the summation error (< K^2 2^-53)
  }
}

// D[t][j] = sum_b sum_i A[b][j][i] BT[t - b][i] (BT has zero rows -3 .. -1
// and T .. Tp - 1): a block of 4 wavefronts computes a 128 x 128 tile (each
// wave 64 rows x 64 columns, 2 x 2 MFMA 32x32x32 i8 tiles) over the four
// planes in turn; rows, columns and the reduction are padded, so there are
// no edges. One int32 plane out instead of four.
// MFMA operands: lane l supplies row (or column) l % 32 and 16 reduction
// elements, the same 16 for the A and the B operand (here 16 (l / 32) ..
// 16 (l / 32) + 15 of the 32); accumulator element e of lane l is row
// 8 (e / 4) + 4 (l / 32) + e % 4, column l % 32.
__global__ void __launch_bounds__(256) k_crt_gemm(const CrtArgs g) {
  // The block's 128 rows and 128 columns of each 32-wide reduction step are
  // staged once in LDS (each thread copies 16 bytes of each), double
  // buffered, and the four waves read their MFMA operands from there: half
  // the L2 traffic of every wave loading its own. Rows are LS bytes apart in
  // LDS (32 bytes of data, 16 of padding) to spread the 16-byte reads over
  // the banks.
  constexpr int LS = 48;
  __shared__ __attribute__((aligned(16))) int8_t la[2][128 * LS];
  __shared__ __attribute__((aligned(16))) int8_t lb[2][128 * LS];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int brow = blockIdx.y * 128, bcol = blockIdx.x * 128;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;  // the wave's 64 x 64 of the block's tile
  const int kl = (lane >> 5) * 16, rl = lane & 31;
  const int lr = tid >> 1, lh = (tid & 1) * 16;  // this thread's 16 staged bytes: row lr, bytes lh ..
  const int nk = g.Kp / 32, nsteps = 4 * nk;   // (plane, 32-wide slice) steps
  // row lr of the block in plane b (A) and digit row bcol + lr - b of the table (B; rows -3 .. -1 are zero)
  auto ga = [&](int st) -> const v4i* {
    const int b = st / nk, k0 = (st - b * nk) * 32;
    return reinterpret_cast<const v4i*>(g.A + (size_t)b * g.npad * g.Kp + (size_t)(brow + lr) * g.Kp + k0 + lh);
  };
  auto gb = [&](int st) -> const v4i* {
    const int b = st / nk, k0 = (st - b * nk) * 32;
    return reinterpret_cast<const v4i*>(g.bt + (ptrdiff_t)(bcol + lr - b) * g.Kp + k0 + lh);
  };
  v4i ra = *ga(0), rb = *gb(0);
  *reinterpret_cast<v4i*>(&la[0][lr * LS + lh]) = ra;
  *reinterpret_cast<v4i*>(&lb[0][lr * LS + lh]) = rb;
  __syncthreads();
  v16i c00 = {0}, c01 = {0}, c10 = {0}, c11 = {0};
  for (int st = 0; st < nsteps; ++st) {
    const int cur = st & 1;
    const bool next = st + 1 < nsteps;
    if (next) {  // the next step's global loads are in flight during this step's MFMAs
      ra = *ga(st + 1);
      rb = *gb(st + 1);
    }
    const v4i a0 = *reinterpret_cast<const v4i*>(&la[cur][(wr + rl) * LS + kl]);
    const v4i a1 = *reinterpret_cast<const v4i*>(&la[cur][(wr + 32 + rl) * LS + kl]);
    const v4i b0 = *reinterpret_cast<const v4i*>(&lb[cur][(wc + rl) * LS + kl]);
    const v4i b1 = *reinterpret_cast<const v4i*>(&lb[cur][(wc + 32 + rl) * LS + kl]);
    c00 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, b0, c00, 0, 0, 0);
    c01 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, b1, c01, 0, 0, 0);
    c10 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b0, c10, 0, 0, 0);
    c11 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b1, c11, 0, 0, 0);
    if (next) {  // the other buffer was last read in the previous step, before its barrier
      *reinterpret_cast<v4i*>(&la[cur ^ 1][lr * LS + lh]) = ra;
      *reinterpret_cast<v4i*>(&lb[cur ^ 1][lr * LS + lh]) = rb;
    }
    __syncthreads();
  }
  const int row0 = brow + wr, col0 = bcol + wc;
  int32_t* C = g.C;
  const int rq = (lane >> 5) * 4;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int r = (e >> 2) * 8 + rq + (e & 3);
    C[(size_t)(col0 + rl) * g.npad + row0 + r] = c00[e];
    C[(size_t)(col0 + 32 + rl) * g.npad + row0 + r] = c01[e];
    C[(size_t)(col0 + rl) * g.npad + row0 + 32 + r] = c10[e];
    C[(size_t)(col0 + 32 + rl) * g.npad + row0 + 32 + r] = c11[e];
  }
}

// s (S - k M) into NL = L + 2 limbs of one row (limb l from D[4l .. 4l+3]),
// eight limbs per round so that the 32 loads of a round are in flight
// together; returns the final carry (the sign word: 0 or -1).
__device__ __forceinline__ int64_t fold_limbs(const CrtArgs& g, int j, int64_t kq, int64_t s, uint32_t* out,
                                              int NL) {
  int64_t carry = 0;
  for (int l0 = 0; l0 < NL; l0 += 8) {
    int32_t cv[8][4];
#pragma unroll
    for (int lq = 0; lq < 8; ++lq)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        // every load issued (a clamped row, then a select), so the 32 are in
        // flight together; a load under the bound's branch waited for each
        const int t = 4 * (l0 + lq) + q;
        const int32_t x = g.C[(size_t)(t < g.Tp ? t : g.Tp - 1) * g.npad + j];
        cv[lq][q] = t < g.Tp ? x : 0;
      }
#pragma unroll
    for (int lq = 0; lq < 8; ++lq) {
      const int l = l0 + lq;
      if (l >= NL) break;
      int64_t v = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) v += (int64_t)cv[lq][q] * ((int64_t)1 << (8 * q));
      v = s * (v - kq * (int64_t)g.m[l]) + carry;  // |v| < 2^53
      const uint32_t limb = (uint32_t)v;
      carry = (v - (int64_t)limb) >> 32;
      out[l] = limb;
    }
  }
  return carry;
}

// One lane per row. Away from the edges of (-1/2, 1/2), the FP64 estimate of
// x / M fixes the sign of x = S - k M and |x| < (M - 1) / 2, so one pass of
// -(S - k M) or S - k M gives the magnitude (checked: its final carry is
// 0). Otherwise (or if the check fails) the exact path: S - k M in two's
// complement, at most one +-M against (M - 1) / 2, then the magnitude.
__global__ void __launch_bounds__(256) k_crt_fold(const CrtArgs g) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= g.n) return;
  const int NL = g.L + 2;
  const int64_t kq = g.kq[j];
  const double d = g.fd[j];
  uint32_t* out = g.limbs + (size_t)g.rows[j] * g.res_stride;
  const bool fast = __builtin_fabs(d) > 1e-6 && __builtin_fabs(d) < 0.5 - 1e-6;
  const int64_t s = (fast && d < 0.0) ? -1 : 1;
  int64_t sw = fold_limbs(g, j, kq, s, out, NL);
  bool neg;
  if (fast && sw == 0) {
    neg = s < 0;
  } else {
    if (s < 0) sw = fold_limbs(g, j, kq, 1, out, NL);
    // x = sum out[l] 2^(32 l) + sw 2^(32 NL), sw in {-1, 0}; |x| < 3M/2
    int fix = 0;  // -1: x -= M, +1: x += M
    if (sw == 0) {
      int cmp = 0;  // x against H = (M - 1) / 2, from the top limb
      for (int l = NL - 1; l >= 0 && cmp == 0; --l) {
        const uint32_t a = out[l], h = g.h[l];
        cmp = a > h ? 1 : (a < h ? -1 : 0);
      }
      if (cmp > 0) fix = -1;
    } else {
      int64_t c = 0;  // x + H: its sign word is sw + c
      for (int l = 0; l < NL; ++l) c = ((int64_t)out[l] + (int64_t)g.h[l] + c) >> 32;
      if (sw + c < 0) fix = 1;
    }
    if (fix != 0) {
      int64_t c = 0;
      for (int l = 0; l < NL; ++l) {
        const int64_t v = (int64_t)out[l] + (int64_t)fix * (int64_t)g.m[l] + c;
        const uint32_t limb = (uint32_t)v;
        c = (v - (int64_t)limb) >> 32;
        out[l] = limb;
      }
      sw += c;
    }
    neg = sw < 0;
    if (neg) {  // magnitude: two's-complement negation
      uint64_t c = 1;
      for (int l = 0; l < NL; ++l) {
        const uint64_t v = (uint64_t)(uint32_t)~out[l] + c;
        out[l] = (uint32_t)v;
        c = v >> 32;
      }
    }
  }
  for (uint64_t l = NL; l < g.res_stride; ++l) out[l] = 0u;
  g.neg[g.rows[j]] = (uint8_t)(neg ? 1 : 0);
}

hipError_t launch_crt_matrix(const CrtArgs& g, hipStream_t s) {
  if (g.n <= 0) return hipSuccess;
  if (g.npad % 128 || g.Tp % 128 || g.Kp % 32 || g.npad < g.n || g.Kp < g.K || g.Tp < g.T ||
      (size_t)g.L + 2 > g.res_stride || g.Tp < g.T + 3)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_crt_digits, dim3(g.npad), dim3(256), 0, s, g);
  hipLaunchKernelGGL(k_crt_gemm, dim3(g.Tp / 128, g.npad / 128), dim3(256), 0, s, g);
  hipLaunchKernelGGL(k_crt_fold, dim3((g.n + 255) / 256), dim3(256), 0, s, g);
  return hipGetLastError();
}

}  // namespace drx
