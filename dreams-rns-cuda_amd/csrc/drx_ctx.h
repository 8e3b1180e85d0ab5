// Internal context / plan structures of libdreams_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dreams_hip.h"
#include "drx_kernels.h"
#include "drx_pack.h"
#include "drx_types.h"

namespace drx {

// Growable device buffer.
struct DBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = bytes < 256 ? 256 : bytes + bytes / 4;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      want = bytes;
      e = hipMalloc(&p, want);
      if (e != hipSuccess) {
        p = nullptr;
        return e;
      }
    }
    cap = want;
    return hipSuccess;
  }
  // Like ensure, but an outgrown buffer is handed to `retired` instead of
  // freed: launch arguments captured earlier (plans, graphs) stay valid.
  hipError_t ensure_keep(size_t bytes, std::vector<void*>& retired) {
    if (bytes <= cap) return hipSuccess;
    if (p) retired.push_back(p);
    p = nullptr;
    cap = 0;
    return ensure(bytes);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

struct HostProgram {
  PackedProgram pk;
  DevProgram dev;
  int mode = 0;  // DRX_WALK_MATRIX or DRX_WALK_VECTOR packing
  int fd = 0;  // runs on the forward-difference kernel
  // executed work per step, for the roofline: modular multiplies / adds
  int mm_ops = 0, ma_ops = 0;
  // the same when only rows 0 and m - 1 of the product are carried
  int mm_ops_rows = 0, ma_ops_rows = 0;
};

struct Timing {
  double walk_ms = 0, combine_ms = 0, shadow_ms = 0, tail_ms = 0, total_ms = 0;
  int walk_launches = 0;
};

// Matrix-core CRT tables of the first K primes (drx_crt.h), built once per K.
struct CrtTable {
  int K = 0, Kp = 0, T = 0, Tp = 0, L = 0;
  DBuf d_bt, d_w, d_m, d_h;
};

}  // namespace drx

struct drx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;   // residue walk, combine, tails
  hipStream_t stream2 = nullptr;  // FP64 M_f pass, forked from `stream`
  hipStream_t stream3 = nullptr;  // FP64 shadow, window by window behind stream2, joined into `stream`
  // the same three streams with CU masks: the FP64 ones on `masked_cus`
  // CUs spread over the chip, the first on the others (made on first use)
  hipStream_t mstream[3] = {nullptr, nullptr, nullptr};
  int masked_cus = 0;
  std::string err;
  // primes
  std::vector<uint32_t> primes;
  drx::DBuf d_prime_p, d_prime_pinv, d_prime_e, d_prime_u;
  drx::DBuf d_invs;  // [K][kFdInv + 1] balanced 1/t mod p_k (forward-difference table jumps)
  std::vector<uint32_t> inv_table;  // [kinv][kinv]: inv_table[j*kinv+i] = p_j^{-1} mod p_i (kinv <= 256)
  int kinv = 0;
  drx::DBuf d_inv;
  // wide Garner: w_i = (p_0 ... p_{i-1})^{-1} mod p_i for the first garner_w.size() primes
  std::vector<double> garner_w;
  drx::DBuf d_garner_w;
  // leading primes in (2^30, 2^31): walks take at most this many (the CRT takes any odd primes)
  int kwalk = 0;
  // device address of the allocation-free CRT's staging array on this device (drx_noalloc.hip)
  void* scratch = nullptr;
  // programs
  std::vector<drx::HostProgram> progs;
  std::vector<uint32_t> h_code;
  std::vector<double> h_cres, h_cflt;
  drx::DBuf d_code, d_cres, d_cflt, d_progs;
  bool progs_dirty = true;
  // scratch for standalone CRT / top-k
  drx::DBuf d_tmp0, d_tmp1, d_tmp2, d_tmp3;
  // matrix-core CRT: tables per K, scratch of the standalone CRT
  std::vector<std::unique_ptr<drx::CrtTable>> crt_tables;
  drx::DBuf d_crt_rows, d_crt_a, d_crt_c, d_crt_k;
  drx::Timing timing;
  hipEvent_t ev[24] = {};
  // Bumped when the prime table changes other than by extension or the
  // programs are released: plans and server requests of an older epoch
  // refer to tables that no longer exist and are refused.
  uint64_t epoch = 0;
  // Outgrown shared device buffers (primes, programs, Garner table), kept
  // until the context ends because plans built earlier still point at them.
  std::vector<void*> retired;
  std::mutex mu;
};

// Shared host helpers (drx_api.cpp).
namespace drx {

// Events of the context: [0..4] timing marks of the last timed launch
// sequence, [5] early output copies done, [6] fork, [7] join back, [8] combined
// residues final, [9] CRT limbs final, [10..11] copy calibration, [12..] one
// per M_f window (the shadow of window c waits for it).
enum { kEvCopied = 5, kEvFork = 6, kEvJoin = 7, kEvOut = 8, kEvLimbs = 9, kEvWin = 12, kMaxWin = 8 };

// One group of a launch: walks of one matrix order and walk-kernel family.
struct Stage {
  int m = 0, kind = 0;             // WalkKind
  int lclass = 0;                  // lds_class of its walks
  int task_begin = 0, task_end = 0;
  int span_begin = 0, span_end = 0;
  int seed_begin = 0, seed_end = 0;    // seed tasks (forward-difference stages)
  int range_begin = 0, range_end = 0;  // M_f blocks, ordered by window
  std::vector<int> win_range;          // [nwin + 1]: ranges of window c are win_range[c] .. win_range[c + 1]
  int threads = 0, seed_kp = 0;    // walk workgroup; seed primes per point group (pow2 >= every K)
  int persist = 0;                 // > 0: persistent walk kernel of this many workgroups (task queue)
  int wpg = 1;                     // walks (tasks) per workgroup; > 1: groups group_begin .. group_end
  int group_begin = 0, group_end = 0;  // of the plan's group table, wpg task indices each
  size_t lds = 0, seed_lds = 0, mf_lds = 0;
};

// Launch the walk pipeline of `stages` with the argument block `a`: on
// stream2 the FP64 M_f pass and the shadow (with delta_float when
// `finalize`), on stream the Newton-table seeds (tasks from `seed_tasks`),
// the walk kernels and the ordered combine. join_stages makes stream wait for
// stream2 (work on stream before it overlaps the shadow chain). With
// `timed`, events 0..3 mark walk start, walk end, combine end, join.
// The FP64 chain runs in `nwin` windows of `win` steps: the M_f pass of
// window c on stream2 (with several windows, the first alone on stream ahead
// of the seeds), the shadow of window c on stream3 behind it.
// With `mf_first` the whole M_f pass runs on `stream` before the walk
// kernels, with the device to itself, and the shadow windows follow it on
// stream3 while the walk runs: for batches whose walk work dwarfs the M_f
// pass (interpreted walks), where a pass interleaved with the walk would be
// starved of workgroup slots and end long after it.
int launch_stages(drx_ctx* ctx, const KArgs& a, const std::vector<Stage>& stages, const DevTask* seed_tasks,
                  int finalize, bool timed, int64_t win, int nwin, bool mf_first = false);
int join_stages(drx_ctx* ctx, bool timed);
// Window length (a multiple of kMfRange) and count for walks of up to
// `max_steps` steps: at most kMaxWin windows of at least two M_f ranges.
void shadow_windows(int64_t max_steps, int64_t* win, int* nwin);
// Orders the M_f ranges of every stage by window (a range belongs to the
// window of its first step) and fills Stage::win_range.
void order_ranges(std::vector<Stage>& stages, std::vector<DevRange>& ranges, int64_t win, int nwin);
// LDS bytes of the kernels of one walk (program `pg`, K primes) in a stage.
size_t walk_lds_bytes(const DevProgram& pg, int kind, int m, uint32_t K, int threads, int wpg = 1,
                      int slice = kSliceK);
size_t seed_lds_bytes(const DevProgram& pg, int kp, int seed_threads);
size_t mf_lds_bytes(const DevProgram& pg);
// Walk-kernel family of program `hp` (forward differences unless interp_only).
int walk_kind(const HostProgram& hp, bool interp_only);
// Task shape of a walk of this kind (the forward-difference kernels' case).
int walk_shape(const HostProgram& hp, int kind);
// Last-column degree of a companion-shaped polynomial program of order 3 or
// 4 (the companion forward-difference kernel), else -1.
int comp_shape(const DevProgram& d);
// Interpreted walks are grouped by their spill rows (0, 1-8, 9-16, more),
// so that a program that spills much does not set the LDS size (and the
// occupancy) of every walk of its order.
int lds_class(const HostProgram& hp, int kind);
// Joins adjacent stages of different LDS classes whose merged LDS size keeps
// the occupancy of the smaller.
void merge_lds_classes(std::vector<Stage>& stages);
// LDS one workgroup may allocate on the context's device (64 KiB if unknown).
size_t device_lds_limit(drx_ctx* ctx);
// Task shape for WALK_FD_PCF: 8 * deg b + deg a.
int pcf_shape(const HostProgram& hp);
int fail(drx_ctx* ctx, int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
bool is_prime_u32(uint32_t n);
double now_s();
int upload_programs(drx_ctx* ctx);
int ensure_inv_table(drx_ctx* ctx, int k);
// The wide Garner table for the first k primes (extended on demand).
int ensure_garner_w(drx_ctx* ctx, int k);
// Looks up the allocation-free CRT's device scratch for ctx->device (the
// device must be current); called once by drx_ctx_create.
int noalloc_init(drx_ctx* ctx);
// The matrix-core CRT table of the first K primes (built on first use).
int ensure_crt_table(drx_ctx* ctx, int K, const CrtTable** out);

inline uint32_t mod_i64(int64_t c, uint32_t p) {
  int64_t r = c % (int64_t)p;
  return (uint32_t)(r < 0 ? r + p : r);
}
inline uint32_t mod_i128(__int128 c, uint32_t p) {
  __int128 r = c % (__int128)p;
  return (uint32_t)(r < 0 ? r + p : r);
}
inline uint32_t powmod(uint32_t a, uint64_t e, uint32_t p) {
  uint64_t r = 1 % p, x = a % p;
  while (e) {
    if (e & 1) r = r * x % p;
    x = x * x % p;
    e >>= 1;
  }
  return (uint32_t)r;
}
inline double balanced(uint32_t r, uint32_t p) { return r > p / 2 ? (double)r - (double)p : (double)r; }
inline uint32_t pow2_ceil(uint32_t k) {
  uint32_t kp = 1;
  while (kp < k) kp <<= 1;
  return kp;
}

// Residue and FP64 tables of one walk's axes: axis i at step s is
// (c0 + s*c1) * inv(den) with c0 = sn*td, c1 = tn*sd, den = sd*td.
// axr gets [dim][3][K] balanced residues, axi [dim][2] int64, axd [dim].
int axis_tables(drx_ctx* ctx, const int64_t* q, int dim, int K, int64_t depth, double* axr,
                int64_t* axi, double* axd);
// Initial matrix of a walk: res [m*m][K] (exact residues as doubles), f [m*m].
int init_tables(drx_ctx* ctx, const drx_walk& w, int m, const uint32_t* init_res_pool,
                const double* init_f_pool, double* res, double* f);
}  // namespace drx

#define DRX_HIPCHK(ctx, expr)                                                                   \
  do {                                                                                          \
    hipError_t _e = (expr);                                                                     \
    if (_e != hipSuccess) return ::drx::fail((ctx), DRX_EHIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

