// Kernel argument blocks and launchers (internal to libdreams_hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "drx_types.h"

namespace drx {

// Pointers into the context's device pools for one batch launch.
struct KArgs {
  const DevTask* tasks;
  const DevWalk* walks;
  const DevProgram* progs;
  const DevSpan* spans;
  const DevRange* ranges;  // M_f blocks
  const uint32_t* code;
  const double* cres;
  const double* cflt;
  const double* prime_p;
  const double* prime_pinv;
  const uint32_t* prime_e;
  const double* invs;     // [K][kFdInv + 1]: 1/t mod p_k (balanced), t >= 1
  const double* axr;
  const int64_t* axi;
  const double* axd;
  const double* res_in;  // walk state before this launch: residues
  const double* fst_in;  // walk state before this launch: FP64 shadow state
  double* res;        // walk state after: residues (may alias res_in)
  double* fst;        // walk state after: FP64 shadow state (may alias fst_in)
  double* cp;         // chunk products
  double* mf;         // shadow matrices M_f(s)
  double* fdt;        // forward-difference Newton tables at step 0
  double* fout;       // per walk: p_f, q_f, log_scale, delta, log_q, (pad)
  int64_t* rout;      // [out_row][2][out_ld]: p residues then q residues
  int64_t* pout;      // optional [out_row][16][out_ld] full state residues
  double* fcompact;   // optional [out_row][6]: copy of fout of the finalized walks
  double* ffull;      // optional [out_row][16]: final P_f
  uint64_t out_ld;
  const DevVGroup* vgroups;  // vector walks: residue groups
  const DevVGroup* fgroups;  // vector walks: FP64 groups (one lane per walk)
  const int32_t* vorder;     // vector walks: walk indices grouped
  uint32_t* queue;           // persistent walk kernels: one task counter per stage
  // walks per workgroup: with wpg > 1 a walk kernel's unit of work is a
  // group of wpg tasks, gtask[unit * wpg + g] (-1: an empty slot), lane
  // group g of the workgroup on task g; the plan groups tasks that share the
  // program (and prime slice) for the interpreter, the entry kinds and shape
  // for the forward-difference kernels, and the prime stripe width
  const int32_t* gtask;
  int32_t wpg;
  double* axr_w;             // server admission: writable views of axr / axi / axd
  int64_t* axi_w;
  double* axd_w;
};

// Words of padding past the end of the code pool (both codes are whole
// groups of kResGroup instructions, so no read passes the end; the padding
// is a margin).
constexpr int kCodeFetch = 8;

// Walk kernel families.
enum WalkKind { WALK_INTERP = 0, WALK_FD = 1, WALK_FD_PCF = 2, WALK_FD_COMP = 3 };
// M_f pass: one lane per step, kMfBlock lanes per block; a range (one block)
// covers kMfRange consecutive steps of one task.
constexpr int kMfBlock = 256;
constexpr int kMfRange = 2 * kMfBlock;  // each lane evaluates 2 steps
// Doubles of slack after the M_f pool: the shadow kernels load their
// register tiles whole, up to two tiles (2 x 8 steps of 4 entries, or 2 x 4
// steps of 16) past a walk's last step, into the next walk's M_f or this
// slack; nothing loaded there is used.
constexpr int kMfSlack = 256;

struct GarnerArgs {
  const int64_t* res;     // rows of canonical residues, row stride res_stride
  uint64_t res_stride;
  const int32_t* rows;    // optional [n] row indices (else rows 0 .. n)
  int n;
  int K;                  // primes per row unless row_k is given
  const int32_t* row_k;   // optional per-row prime count
  const double* prime_p;
  const double* prime_pinv;
  const uint32_t* prime_u;
  const uint32_t* inv_table;  // inv_table[j * inv_ld + i] = p_j^{-1} mod p_i
  uint64_t inv_ld;
  uint32_t* limbs;        // [n][res_stride] little-endian magnitude of the centred value (zero past K)
  uint8_t* neg;           // [n] 1 if the centred value is negative
  const double* w;        // wide Garner: w_i = (p_0 ... p_{i-1})^{-1} mod p_i (canonical)
};

// Wide Garner (rows of any width up to kMaxK): one workgroup per row.
// Workgroup size and LDS bytes for rows of up to kmax primes.
inline int garner_wide_threads(int kmax) { return kmax <= 1024 ? 256 : 1024; }
// Digits (and limbs) per thread: the kernel's template width, 1, 2, 4 or 8.
inline int garner_wide_regs(int kmax) {
  const int nt = garner_wide_threads(kmax);
  int r = 1;
  while (r * nt < kmax) r *= 2;
  return r;
}
inline size_t garner_wide_lds(int kmax) {
  const int nt = garner_wide_threads(kmax);
  return (size_t)8 * nt * garner_wide_regs(kmax) + (size_t)8 * nt + 64;
}


// Walk kernel of one stage over a.tasks[0 .. ntask): one workgroup per task,
// or with persist_blocks > 0 a persistent grid of that many workgroups
// (at most ntask) taking tasks from the counter `queue` (cleared before).
hipError_t launch_walk(int m, int kind, const KArgs& a, int ntask, int threads, size_t lds, hipStream_t s,
                       int persist_blocks = 0, uint32_t* queue = nullptr);
hipError_t launch_fill_u32(uint32_t* p, int n, uint32_t v, hipStream_t s);
// Vector walks: residues (one block per residue group of `threads` lanes) and
// the FP64 shadow (one block per FP64 group, one lane per walk).
hipError_t launch_vec_walk(const KArgs& a, int ngroup, int threads, size_t lds, hipStream_t s);
hipError_t launch_vec_fp(const KArgs& a, int ngroup, int threads, size_t lds, hipStream_t s);
// Resident workgroups per CU of the walk kernel of order m and family kind
// (its grid or its persistent variant) at this workgroup size and LDS; -1
// when unknown.
int walk_blocks_per_cu(int m, int kind, int threads, size_t lds, bool persistent = false);
// Newton tables of forward-difference walks: one block per seed task (a
// prime slice of a walk, a.tasks) of `threads` lanes, `kp` (power of two >=
// every slice's prime count) per point group; threads % kp == 0.
hipError_t launch_fd_seed(int m, const KArgs& a, int nseed, int kp, int threads, size_t lds, hipStream_t s);
// Seed block for primes padded to kp: point groups of kp lanes, at most 256
// lanes (the kernel needs over 128 VGPRs, so a CU holds 8 of its waves: two
// blocks of 256 lanes overlap each other's latency where one block of 512
// would leave the CU idle between blocks; 3.3 ms against 5.5 ms for the
// 4x4 seeds of 8256 walks at K = 1000).
// Seed launches of fewer lanes than this split a task's points over
// workgroups (a small batch's tasks alone leave most of the device idle).
constexpr size_t kSeedSplitLanes = (size_t)1 << 20;
inline int fd_seed_threads(int m, int kp) {
  const int pts = fd_dmax(m) + 1;
  int j = (kp <= 64 ? 64 : 256) / kp;
  if (j < 1) j = 1;
  if (j > pts) j = pts;
  return j * kp;
}
// FP64 M_f(s) over `nrange` blocks of a.ranges.
hipError_t launch_mf(int m, const KArgs& a, int nrange, size_t lds, hipStream_t s);
hipError_t launch_combine(int m, const KArgs& a, int nspan, hipStream_t s);
// FP64 shadow over steps [w0, w1) of every span's walk (w0 == 0 starts from
// fst_in, later windows continue from fst); delta_float when `finalize`.
hipError_t launch_shadow(int m, const KArgs& a, int nspan, int finalize, int64_t w0, int64_t w1, hipStream_t s);
hipError_t launch_finalize(const KArgs& a, int nspan, int full, hipStream_t s);
// Server admission: initial residues (canonical, as the host tables), FP64
// state and axis tables of every admitted walk, written into its slot
// (a.walks[rec.slot] gives the offsets). One block per record.
hipError_t launch_admit(const KArgs& a, const DevAdmit* recs, int n, hipStream_t s);
hipError_t launch_garner(const GarnerArgs& g, hipStream_t s);
// Rows of up to g.K primes (g.K <= kMaxK), one workgroup per row; needs g.w.
hipError_t launch_garner_wide(const GarnerArgs& g, hipStream_t s);
// Top-k walk indices by value (larger first, NaN last, ties to the lower
// index). `ws` holds topk_workspace_bytes(n, k) bytes (unused when 0).
struct TopCand {
  uint64_t key;
  uint32_t idx, pad;
};
size_t topk_workspace_bytes(int n, int k);
hipError_t launch_topk(const double* d, int stride, int n, int k, int32_t* out, void* ws, hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, size_t n16, hipStream_t s);
// Issue-rate probe (kRateChains chains per lane, `iters` rounds) of op 0..3.
constexpr int kRateChains = 8;
hipError_t launch_rate(int op, int blocks, int iters, double* out, hipStream_t s);

}  // namespace drx
