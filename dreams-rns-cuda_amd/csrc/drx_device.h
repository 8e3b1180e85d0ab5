// Device-side arithmetic and bytecode encoding for the walk kernels (gfx950).
//
// Residue arithmetic runs on the FP64 pipes, which are full rate on CDNA4:
// a residue is an integer held exactly in a double, kept in a "balanced",
// lazily reduced range. For a prime p in [2^30, 2^31) with pinv = fl(1/p):
//
//   mm(a, b):  h = fl(a*b); l = fma(a, b, -h)        (a*b = h + l exactly)
//              q = rint(h * pinv); r = fma(-q, p, h) + l
//
// r == a*b - q*p exactly whenever |a|, |b| <= 2^40 (then |h - q p| < 2^53 and
// the fma is exact), and |r| <= p/2 + 1.5 * 2^-52 * |a b| < 2^31. Additions
// are a single fadd with no reduction; the host-side packer tracks magnitude
// bounds and inserts RED (in-place reduction) before any value could exceed
// 2^40. Canonical residues in [0, p) are produced only at the very end.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace drx {

// ------------------------------------------------------- residue encoding
// The residue code is one 32-bit word per instruction over one bank B of 32
// doubles per lane in VGPRs: B[0, 16) are temps, B[16 + j] is slot j of the
// inputs C (C[15] = step, C[14 - i] = axis i, C[0 .. 15 - dim) = the first
// constant slots). Both operands come from anywhere in B and every result
// goes to a temp, so an arithmetic instruction costs two indexed register
// reads and one indexed write (one index window each) and one or two
// uniform branches.
//   [3:0] op   [8:4] x   [13:9] y   [18:14] dst   [31:19] aux
// aux is an LDS row for LDM / STM / LDC / LDV / STE, or the LIN signs: sx in
// [20:19] and sy in [22:21], each a 2-bit two's complement number in
// {-1, 0, +1}; LIN computes sy * y + sx * x (SUB, NEG, zero). A matrix entry
// (in a vector walk: an entry of the new state vector) leaves through STE,
// which stores x to the lane's LDS row of that entry. Every instruction
// writes dst; STM, STE and NOP, which compute nothing, carry dst = x and
// write x back.
enum ResOp : uint32_t {
  RO_MUL = 0,    // x * y
  RO_ADD = 1,    // x + y
  RO_LIN = 2,    // sy * y + sx * x
  RO_MULNZ = 3,  // x * (y, or 1 when y = 0 mod p)
  RO_ZSEL = 4,   // x, or 0 when y = 0 mod p
  RO_INV = 5,    // x^(p-2)
  RO_RED = 6,    // x mod p (balanced)
  RO_NZ1 = 7,    // x, or 1 when x = 0 mod p
  RO_LDM = 8,    // the lane's spill row
  RO_LDC = 9,    // LDS constant row (constant slot 15 - dim + row)
  RO_LDV = 10,   // the lane's state row (vector walks)
  RO_STM = 11,   // spill row <- x (dst = x)
  RO_STE = 12,   // entry row <- x (dst = x)
  RO_NOP = 15,   // nothing: B[x] <- B[x] (pads the code to whole groups)
};
// The FP64 code of the M_f pass (and of vector walks' shadow) uses the same
// encoding, bank layout and register allocation with FP64 values in bank B
// (the constant slots hold the constants' FP64 values): MUL, ADD and LIN
// (SUB) are the FP64 operations, the LDS and no-op codes are shared, and
// these replace the residue-only ones:
enum FpOp : uint32_t {
  FO_NEG = 3,   // -x
  FO_CUB = 4,   // x^3, correctly rounded
  FO_INV = 5,   // 1/x, or 0 when |x| <= 1e-300
  FO_MOV = 6,   // x
  FO_ZERO = 7,  // +0
};
constexpr int kResTemps = 16;    // temps B[0, 16)
constexpr int kResCBase = 16;    // input slot j at B[16 + j]
constexpr int kResStep = 15;     // input slot of the step; axis i at slot 14 - i
__host__ __device__ constexpr uint32_t enc_res(uint32_t op, uint32_t x, uint32_t y, uint32_t d, uint32_t aux) {
  return op | (x << 4) | (y << 9) | (d << 14) | (aux << 19);
}
// LIN signs (each in {-1, 0, +1}) as the aux field.
__host__ __device__ constexpr uint32_t lin_signs(int sx, int sy) {
  return ((uint32_t)sx & 3u) | (((uint32_t)sy & 3u) << 2);
}
// The packer emits one word per instruction; in the device's code pool
// each is stored unpacked as kResWords words {x, y, dst, op | aux << 4}, so
// the interpreter takes the indices straight from the scalar registers the
// code is loaded into, with no field extraction.
constexpr int kResWords = 4;
struct ResWords {
  uint32_t w[kResWords];
};
__host__ __device__ constexpr ResWords unpack_res(uint32_t w) {
  return ResWords{{(w >> 4) & 31u, (w >> 9) & 31u, (w >> 14) & 31u, (w & 15u) | ((w >> 19) << 4)}};
}
// The residue code is executed in groups of kResGroup instructions read by
// two scalar loads; the packer pads it with no-ops.
constexpr int kResGroup = 8;
constexpr uint32_t kResNop = RO_NOP;
constexpr uint32_t kMaxRow = (1u << 13) - 1;

// ------------------------------------------------ value operations
// Operations of the packer's value graph (and of its FP64 code list); the
// device codes above are generated from them.
enum DevOp : uint32_t {
  D_ADD = 1,   // a + b
  D_SUB = 2,   // a - b
  D_MUL = 3,   // a * b
  D_NEG = 4,   // -a
  D_SQR = 5,   // a^2
  D_CUB = 6,   // a^3
  D_INV = 7,   // a^(p-2) mod p          | float: 1/a if |a| > 1e-300 else 0
  D_RED = 8,   // a mod p (balanced)     | float: a
  D_MOV = 9,   // a
  D_NZ1 = 10,  // a, or 1 when a = 0 mod p (residue code only)
  D_ZSEL = 11, // a, or 0 when b = 0 mod p (residue code only)
  D_MULNZ = 12,  // a * (b, or 1 when b = 0 mod p) (residue code only)
};

// Matrix entry kinds (2 bits per entry, entry e = row*m + col).
enum EntryKind : uint32_t { E_ZERO = 0, E_ONE = 1, E_GEN = 2 };

// ------------------------------------------------------------ modular FP64
__device__ __forceinline__ double mm(double a, double b, double p, double pinv) {
  const double h = a * b;
  const double l = __builtin_fma(a, b, -h);
  const double q = __builtin_rint(h * pinv);
  const double r = __builtin_fma(-q, p, h);
  return r + l;
}

// In-place balanced reduction, exact for |x| < 2^52.
__device__ __forceinline__ double red(double x, double p, double pinv) {
  const double q = __builtin_rint(x * pinv);
  return __builtin_fma(-q, p, x);
}

// Canonical residue in [0, p) of an exact integer |x| < 2^52.
__device__ __forceinline__ double canon(double x, double p, double pinv) {
  double r = __builtin_fma(-__builtin_floor(x * pinv), p, x);
  r = (r < 0.0) ? r + p : r;
  r = (r >= p) ? r - p : r;
  return r;
}

// a^(p-2) mod p, left to right over the 31 bits of e = p - 2 (bit 30 is
// always set for p in [2^30, 2^31)). Fermat gives 0 for a == 0 mod p, the
// reference's INV(0) = 0. Bit by bit: 60 modular multiplies, few registers.
__device__ __forceinline__ double inv_fermat_bits(double a, double p, double pinv, uint32_t e) {
  double r = a;  // bit 30
#pragma unroll 1
  for (int bit = 29; bit >= 0; --bit) {
    r = mm(r, r, p, pinv);
    const double t = mm(r, a, p, pinv);
    r = ((e >> bit) & 1u) ? t : r;
  }
  return r;
}
// The same power in 2-bit windows: two squarings, then one multiply by a,
// a^2 or a^3 as the window's digit says (47 modular multiplies; every lane
// has its own e, so the window schedule is the same for all). The residue
// class is the same as bit by bit, the balanced representative may differ.
__device__ __forceinline__ double inv_fermat(double a, double p, double pinv, uint32_t e) {
  const double a2 = mm(a, a, p, pinv);
  const double a3 = mm(a2, a, p, pinv);
  double r = a;  // bit 30
#pragma unroll 1
  for (int bit = 28; bit >= 0; bit -= 2) {
    r = mm(r, r, p, pinv);
    r = mm(r, r, p, pinv);
    const uint32_t d = (e >> bit) & 3u;
    const double f = d == 1u ? a : (d == 2u ? a2 : a3);
    const double t = mm(r, f, p, pinv);
    r = d ? t : r;
  }
  return r;
}

// --------------------------------------------------------------- float ops
// Correctly rounded x^3: x^2 = h + l and h*x = c + e exactly (FMA), then
// x^3 = c + (e + l*x), whose inner sum is ~2^-52 of c, so the final add
// rounds the exact cube correctly except within ~2^-50 ulp of a tie.
__device__ __forceinline__ double cube_cr(double x) {
  const double h = x * x;
  const double l = __builtin_fma(x, x, -h);
  const double c = h * x;
  const double e = __builtin_fma(h, x, -c);
  const double t = __builtin_fma(l, x, e);
  return c + t;
}

__device__ __forceinline__ double finv(double x) {
  return (__builtin_fabs(x) > 1e-300) ? 1.0 / x : 0.0;
}

}  // namespace drx
