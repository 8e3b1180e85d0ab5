// Matrix-core CRT for wide prime sets (internal to libdreams_hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace drx {

// Rows of at most 256 primes take the one-wavefront Garner kernel. Wider
// rows take the matrix-core CRT when their prime count is at most
// kCrtMatrixMaxK and at least kCrtMatrixMinRows rows share it (the GEMM pads
// rows to 128 and needs a table per prime count); every other wide row takes
// the wide Garner kernel (one workgroup per row, no per-K table).
constexpr int kCrtMatrixMinK = 257;
constexpr int kCrtMatrixMaxK = 1024;
constexpr int kCrtMatrixMinRows = 64;
constexpr int kBtLead = 4;  // zero rows of the digit table before row 0 (the GEMM reads rows t - b >= -3)

struct CrtArgs {
  const int64_t* res;      // canonical residue rows, stride res_stride
  uint64_t res_stride;
  const int32_t* rows;     // [n] row index of compact row j (in res, limbs and neg)
  int n, npad;             // rows, rows padded to 128
  int K, Kp;               // primes, padded to 32 (the reduction)
  int T, Tp;               // base-256 digit columns of the M_i; Tp >= T + 3, padded to 128
  int L;                   // 32-bit limbs of M
  const double* prime_p;
  const double* prime_pinv;
  const double* w;         // [K] (M / p_i)^-1 mod p_i
  const int8_t* bt;        // [Tp][Kp] balanced base-256 digits of M / p_i (column i), zero
                           // rows from T on and three zero rows before row 0 (bt[-3 .. -1])
  const uint32_t* m;       // [L + 2] limbs of M (zeros past L)
  const uint32_t* h;       // [L + 2] limbs of (M - 1) / 2
  int8_t* A;               // [4][npad][Kp] digit planes of the centred c_i
  int32_t* C;              // [Tp][npad] D[t][j]: the coefficients of 256^t in S
  int32_t* kq;             // [npad] rint(sum c_i / p_i)
  double* fd;              // [npad] sum c_i / p_i - kq (the FP64 estimate of x / M)
  uint32_t* limbs;         // out: magnitude limbs, row stride res_stride (zeros past L + 2)
  uint8_t* neg;            // out: 1 if the centred value is negative
};

// Digits, GEMM and fold on stream s (no allocation).
hipError_t launch_crt_matrix(const CrtArgs& g, hipStream_t s);

}  // namespace drx
