// Plain-data records shared by the host packer and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace drx {

constexpr int kMaxM = 4;          // matrix order handled by the kernels
constexpr int kMaxK = 8192;       // primes in one context
constexpr int kMaxDim = 15;       // axes per program (FP64 bank X: axes 0..14, step 15)
constexpr int kMaxConst = 1023;   // constants per program (10-bit field)
constexpr int kMaxCode = 1 << 16; // device words per program
constexpr int kFdMax2 = 6;        // forward-difference degree limit, 2x2 programs
constexpr int kFdMax3 = 3;        // forward-difference degree limit, 3x3 and 4x4 (companion) programs
constexpr int kFdInv = 6;         // small inverses 1/t mod p kept per prime (t <= kFdInv)

// FP64 shadow state per walk: P_f [m*m], then log_scale as a mantissa
// product in [0.5, 1) and a power-of-two exponent (log_scale = ln(mant) +
// exp * ln 2), so it accumulates exactly in step order without a log per step.
__host__ __device__ constexpr int fst_size(int m) { return m * m + 2; }
constexpr int kFstMax = 18;       // fst_size(kMaxM)

// Forward-difference table depth (entries per degree column) for order m.
__host__ __device__ constexpr int fd_dmax(int m) { return m == 2 ? kFdMax2 : kFdMax3; }

// One packed program (device copy; offsets index the context pools).
struct DevProgram {
  uint32_t code_off;    // residue code: first word in the code pool (one word per instruction)
  uint32_t ncode;       // residue code: instructions
  uint32_t fcode_off;   // FP64 code (the M_f pass): first word (even) in the code pool, two per instruction
  uint32_t nfcode;      // FP64 code: instructions
  uint32_t nconst;      // constants referenced by the packed code
  uint32_t kcap;        // primes covered by the constant residue table
  uint64_t cres_off;    // [nconst][kcap] balanced residues (double)
  uint64_t cflt_off;    // [nconst] FP64 values
  uint32_t m, dim;
  uint32_t kinds;       // 2 bits per entry: EntryKind
  uint32_t colred;      // bit j: reduce column j after P*M
  uint32_t const_lds;   // FP64 code: 1 when slots beyond the first 16 are read from bank L
  uint32_t nslots;      // residue code: LDS spill rows per lane
  uint32_t fnslots;     // LDS temp slots (bank S) of the FP64 code
  uint32_t pad;
  uint32_t fd_ok;       // every general entry is a polynomial in the step
  uint32_t fd_degree;   // max degree over entries
  uint64_t fd_deg;      // 4 bits per entry: its degree
};

// One walk's resident state and parameters (device copy).
struct DevWalk {
  int32_t prog;
  int32_t K;
  int32_t m;
  int32_t dim;
  int32_t pe, qe;       // vector walks: state entries of p and q
  uint64_t res_off;     // P residues [m*m][K] (double, balanced); vector walks: v(0) residues [m][K]
  uint64_t fst_off;     // FP64 shadow state, fst_size(m) doubles; vector walks: v_f(0) [m]
  uint64_t fdt_off;     // forward-difference walks: Newton tables at step 0, [m*m][fd_dmax(m)+1][K]
  uint64_t axr_off;     // axis residues [dim][3][K]: c0, c1, c2 (double, balanced)
  uint64_t axi_off;     // axis integers [dim][2]: c0, c1 (int64) for the FP64 axis value
  uint64_t axd_off;     // axis denominators [dim] (double)
  double target;        // L for delta_float
};

// A contiguous step range of one walk processed by one workgroup.
struct DevTask {
  int32_t walk;
  int32_t shape;        // forward-difference walks: companion PCF 8 * deg b + deg a, companion
                        // CMF the last column's degree (| kShapeRows); else 0
  int64_t s_begin;
  int64_t s_end;
  uint64_t mf_off;      // M_f of step s at mf[mf_off + (s - s_begin) * m * m]
  uint64_t cp_off;      // chunk product [m*m][K] in the chunk-product pool
  int32_t k0;           // primes [k0, k0 + kn) of the walk: walks of more than
  int32_t kn;           // kSliceK primes run as several tasks per step range
};

// Companion CMF task whose chunk product is needed in rows 0 and m - 1 only.
constexpr int32_t kShapeRows = 256;

// Widest prime stripe of one walk workgroup; a walk with more primes splits
// each step range into slices of at most this many primes (the primes are
// independent; the slices share the range's chunk product).
constexpr int kSliceK = 256;

// Per-walk span of tasks for the ordered combine / shadow passes.
struct DevSpan {
  int32_t walk;
  int32_t task_begin;   // tasks of this walk are contiguous and in step order
  int32_t task_end;
  int32_t out_row;      // row of this walk in the finalize outputs
};

// Walks of one program, depth and prime count run together by a vector-walk
// kernel: walks vorder[first .. first + n).
struct DevVGroup {
  int32_t first;
  int32_t n;
  int32_t prog;
  int32_t K;
  int64_t depth;
};

// A walk admitted into a server slot: its initial matrix and axis tables
// are built on the device from this record (axis i at step s is
// (c0 + s c1) / den, the combined rational of cmf_walk.py:206-224).
struct DevAdmit {
  int32_t slot;
  int32_t init_kind;   // DRX_INIT_PCF or DRX_INIT_IDENTITY
  int64_t a0;          // DRX_INIT_PCF: A = [[1, a0], [0, 1]]
  int64_t c0[kMaxDim], c1[kMaxDim], den[kMaxDim];
};

// A block of consecutive steps of one task for the FP64 M_f(s) evaluation.
struct DevRange {
  int32_t task;
  int32_t pad;
  int64_t s_begin;
  int64_t s_end;
};

}  // namespace drx
