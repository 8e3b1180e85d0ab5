/* dreams_hip.h -- C ABI of libdreams_hip.so, the MI355X (gfx950) walk engine.
 *
 * The reference (VesterlundCoder/Dreams-RNS-CUDA) exposes this path only as
 * Python functions; it has no FFI. Each entry point below is what a ctypes
 * binding of those functions needs, and names the reference interface it
 * replaces. All functions return DRX_OK (0) or a negative DRX_E* code;
 * drx_last_error() then holds a message. No C++ exception or abort crosses
 * the ABI. Host buffers are caller-owned (C-contiguous); device memory is
 * owned by the context. A context must be used by one thread at a time; the
 * server functions are internally synchronised.
 */
#ifndef DREAMS_HIP_H
#define DREAMS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DRX_ABI_VERSION 5

enum {
  DRX_OK = 0,
  DRX_EINVAL = -1,   /* bad argument (the reference raises ValueError/IndexError) */
  DRX_EHIP = -2,     /* HIP runtime error */
  DRX_ENOMEM = -3,   /* device or host allocation failed */
  DRX_ERANGE = -4,   /* value outside what the engine supports (e.g. m > 4) */
  DRX_ESTATE = -5,   /* call out of sequence (e.g. no primes set) */
  DRX_EAGAIN = -6    /* server: nothing ready / queue full */
};

typedef struct drx_ctx drx_ctx;

/* ---- context --------------------------------------------------------- */
int drx_abi_version(void);
int drx_device_count(int* n);
int drx_ctx_create(int device, drx_ctx** out);
void drx_ctx_destroy(drx_ctx* ctx);
const char* drx_last_error(const drx_ctx* ctx);

/* Prime table shared by every walk of the context; a walk with K primes uses
 * the first K. Replaces generate_rns_primes (python/dreams_rns/runner.py:33-52),
 * whose output the caller passes here (distinct odd primes below 2^31, at most
 * 8192). Walks need their primes in (2^30, 2^31): a walk may use only the
 * leading primes of the table that lie there (DRX_ERANGE otherwise); the CRT
 * takes any of them, as crt_reconstruct (runner.py:57-76) takes any primes.
 * Extending the table keeps existing plans valid; any other change (allowed
 * only while no programs are loaded) makes them fail with DRX_ESTATE. */
int drx_set_primes(drx_ctx* ctx, const int64_t* primes, int K);

/* ---- programs ----------------------------------------------------------
 * Upload one compiled CmfProgram (python/dreams_rns/compiler.py:44-89):
 * `ops` holds n_ops instructions as (op, arg0, arg1, arg2) uint32 quadruples,
 * exactly Instruction.pack (compiler.py:39-41). Constants are given either as
 * int64 values (`consts`, reduced here) or, for constants that do not fit,
 * pre-reduced: `const_res` [n_const][kcap] canonical residues for the first
 * kcap primes plus `const_f` [n_const] FP64 values (Python float(c)). Walks
 * of the program may use the first kcap primes (int64 constants with
 * kcap <= 0: every prime set at upload).
 * Replaces _precompute_const_residues (runner.py:182-195). */
int drx_program_upload(drx_ctx* ctx, const uint32_t* ops, int n_ops, int m, int dim,
                       const int64_t* consts, const uint32_t* const_res, const double* const_f,
                       int n_const, int kcap, int* prog_id);
/* As drx_program_upload, packed for walks of kind `mode`: DRX_WALK_MATRIX
 * (the code computes M(s); m <= 4) or DRX_WALK_VECTOR (the code computes the
 * next state v' = M(s) v of a state-vector walk, cmf_walk.py:435-520;
 * m <= 16). The matrix walk's program and the vector walk's are different
 * uploads of the same bytecode. */
int drx_program_upload_mode(drx_ctx* ctx, const uint32_t* ops, int n_ops, int m, int dim,
                            const int64_t* consts, const uint32_t* const_res, const double* const_f,
                            int n_const, int kcap, int mode, int* prog_id);
int drx_program_info(drx_ctx* ctx, int prog_id, int* n_dev_ops, int* n_const_used,
                     uint32_t* entry_kinds);
/* Drop every program. Plans built before, and server requests submitted
 * before, are refused afterwards (DRX_ESTATE). */
int drx_program_release_all(drx_ctx* ctx);

/* Pack a program without a device (inspection and host-side testing of the
 * packer). words receives the residue code (one word per instruction) and,
 * from the next even word, the FP64 code (two words per instruction), up to
 * max_words in total; const_slots[i] =
 * constant index of device slot i (an index >= n_const names a folded
 * constant); derived receives (op, a, b) per folded constant, its value being
 * op applied to constants a and b (b = -1 for a unary op). meta receives up to
 * max_meta of DRX_PACK_META values: [0] residue-code words, [1] entry kinds
 * (2 bits per entry), [2] columns reduced after P*M, [3] constants in LDS,
 * [4] LDS temp slots of the residue code, [5] constant slots, [6] all entries
 * polynomial (forward differences apply), [7] their maximal degree,
 * [8] FP64-code words, [9] LDS temp slots of the FP64 code, [10] folded
 * constants, [11] INVs sharing one inversion, [12] reference instructions.
 * Returns DRX_EINVAL with the packer's message in err (err_len bytes) for
 * programs the engine cannot run. No counterpart in the reference: its
 * bytecode is interpreted as written (python/dreams_rns/runner.py:217-316). */
#define DRX_PACK_META 13
int drx_program_pack(const uint32_t* ops, int n_ops, int m, int dim, const int64_t* consts, int n_const,
                     int mode, uint32_t* words, int max_words, int* const_slots, int max_slots,
                     int32_t* derived, int max_derived, int64_t* meta, int max_meta, char* err, int err_len);

/* ---- walks ------------------------------------------------------------- */
enum { DRX_INIT_PCF = 0, DRX_INIT_IDENTITY = 1, DRX_INIT_GIVEN = 2 };
enum { DRX_MODE_BATCHED = 0, DRX_MODE_PLAIN = 1, DRX_MODE_PERSISTENT = 2 };

/* One walk: P = P0 * M(0) * M(1) * ... * M(depth-1) mod each of K primes,
 * plus the FP64 shadow. Axis i at step s is (sn_i + s * tn_i / td_i ...) given as
 * four int64 per axis in the axis pool at axis_off: sn, sd, tn, td, i.e. the
 * value (sn*td + s*tn*sd) / (sd*td) (python/dreams_rns/cmf_walk.py:206-224).
 * A PCF walk (runner.py:338-413) is sn = shift_val, tn = direction,
 * sd = td = 1, init_kind DRX_INIT_PCF with a0. */
/* A state-vector walk (mode DRX_WALK_VECTOR, run_cmf_walk_vec,
 * cmf_walk.py:445-531) is v(depth) = M(depth-1) ... M(0) v(0) with p =
 * v[out_p], q = v[out_q]; its program is uploaded with mode DRX_WALK_VECTOR,
 * its initial state given (DRX_INIT_GIVEN: [m][K] residues and [m] FP64
 * entries). A batch holds walks of one mode.
 * K: a matrix walk takes up to 8192 primes (above 256 it runs as prime
 * slices of 256: the primes are independent, the results are those of one
 * walk), a vector walk up to 256, a served walk up to 256. */
enum { DRX_WALK_MATRIX = 0, DRX_WALK_VECTOR = 1 };
typedef struct drx_walk {
  int32_t prog;
  int32_t K;
  int64_t depth;
  int32_t init_kind;
  int32_t mode;        /* DRX_WALK_MATRIX or DRX_WALK_VECTOR */
  int64_t a0;          /* DRX_INIT_PCF: A = [[1, a0], [0, 1]] */
  int64_t init_off;    /* DRX_INIT_GIVEN: first of m*m*K residues ([m*m][K]; vector walks [m][K]) in init_res_pool */
  int64_t init_f_off;  /* DRX_INIT_GIVEN: first of m*m FP64 entries (vector walks m) in init_f_pool */
  int64_t axis_off;    /* index of sn of axis 0 in the axis pool */
  double target;       /* L for delta_float */
  int32_t out_p;       /* vector walks: state entry of p */
  int32_t out_q;       /* vector walks: state entry of q */
} drx_walk;

typedef struct drx_walk_out {
  double p_float;
  double q_float;
  double log_scale;
  double delta_float;
  double log_q;
  double reserved;
} drx_walk_out;

/* Batch I/O. Optional outputs may be NULL. */
typedef struct drx_batch {
  const drx_walk* walks;
  int n;
  const int64_t* axis_pool;
  const uint32_t* init_res_pool;  /* DRX_INIT_GIVEN: canonical residues, see drx_walk */
  const double* init_f_pool;      /* DRX_INIT_GIVEN: FP64 entries, see drx_walk */
  drx_walk_out* out;              /* [n] */
  int64_t* residues;              /* [n][2][res_stride]: p then q residues (canonical) */
  int res_stride;                 /* >= max K */
  int64_t* full;                  /* optional [n][16][res_stride] whole final P (row-major m*m) */
  double* full_f;                 /* optional [n][16] whole final FP64 shadow P_f (row-major m*m) */
  uint32_t* crt_limbs;            /* optional [n][2][res_stride] |centred p|, |centred q| limbs */
  uint8_t* crt_neg;               /* optional [n][2] sign of centred p, q */
  int top_k;                      /* 0: none */
  int32_t* topk_idx;              /* [top_k] walk indices by delta_float desc, ties low index */
} drx_batch;

/* DRX_MODE_PERSISTENT: each stage's walk kernel (one per matrix order and
 * kernel family) runs as a persistent grid of the workgroups the device
 * holds at once, which take the stage's tasks (step ranges of walks, longest
 * first) from a device queue; the whole launch sequence, FP64 shadow streams
 * included, is captured as one hipGraph and replayed per run. Results are
 * those of DRX_MODE_BATCHED. State-vector walks keep their grid kernels. */
typedef struct drx_launch_cfg {
  int mode;         /* DRX_MODE_* */
  int threads;      /* workgroup size, power of two in [32, 256]; 0 = tuned */
  int chunk_steps;  /* max steps of one walk per workgroup task; 0 = tuned */
  int use_graph;    /* capture the launch sequence in a hipGraph and replay it */
  int interp_only;  /* 1: interpret every step even for polynomial programs (no forward differences) */
  int chain_cus;    /* > 0: that many CUs (spread over the chip) run only the FP64 M_f pass and shadow,
                       the walk the others; 0: every stream may use every CU. Not with use_graph. */
  int walks_per_wg; /* > 1: a walk workgroup runs that many walks' step ranges side by side, each on
                       threads / walks_per_wg lanes (fewer sub-chunks per walk); 0 or 1: one; capped
                       where a walk's prime stripe would not fit; 0 = tuned */
  int slice;        /* primes per task (the prime stripe a workgroup's lanes cover): a power of two in
                       [16, 256]; wider walks run as slices of that many primes; 0 = 256 */
} drx_launch_cfg;

/* Run a batch of walks to completion. Replaces run_pcf_walk (runner.py:338),
 * run_cmf_walk (cmf_walk.py:363), run_pcf_walk_batch_gpu / run_cmf_walk_batch_gpu
 * (gpu_walk.py:279, 395) and, with crt_limbs, crt_reconstruct + centered
 * (runner.py:57-81) on the p and q of every walk. */
int drx_walk_batch(drx_ctx* ctx, const drx_batch* b, const drx_launch_cfg* cfg);

/* The same in two phases: drx_plan_create packs and uploads everything
 * (inputs become resident in HBM) and, with use_graph, captures the launch
 * sequence; drx_plan_run executes it (re-runnable: every run starts from the
 * walks' initial matrices) and copies the outputs into the drx_batch buffers
 * given at creation, which must stay valid. When every such buffer is
 * page-locked (drx_host_alloc), the residues and CRT limbs are copied as soon
 * as they are final, while the FP64 shadow still runs, and the run launches on
 * streams even for a captured plan; the results are the same either way. */
typedef struct drx_plan drx_plan;
int drx_plan_create(drx_ctx* ctx, const drx_batch* b, const drx_launch_cfg* cfg, drx_plan** out);
int drx_plan_run(drx_plan* plan);
/* drx_plan_run in two steps (no counterpart in the reference, whose walk
 * functions return host arrays: these let a caller keep the results in HBM
 * and time the device work apart from the copy over PCIe).
 * drx_plan_run_resident runs the whole pipeline and synchronises, leaving
 * every output in the plan's device buffers; drx_plan_fetch copies the last
 * run's outputs into the drx_batch buffers and synchronises. A plan in
 * DRX_MODE_PLAIN copies each walk back as it finishes (that is its point),
 * so for it both behave as drx_plan_run does. */
int drx_plan_run_resident(drx_plan* plan);
int drx_plan_fetch(drx_plan* plan);
/* Work of the plan: residue-steps (sum depth*K) and the modular operation
 * counts summed over residue-steps: [0] residue-steps, [1] modular multiplies,
 * [2] modular additions, [3] walk-kernel tasks. */
int drx_plan_stats(const drx_plan* plan, double* s4);
void drx_plan_destroy(drx_plan* plan);

/* Autotuner: picks the workgroup size (threads; with K primes per sub-chunk
 * this sets the sub-chunks, i.e. step ranges, per workgroup) and the chunk
 * length (steps per workgroup task, i.e. tasks per walk) for this batch.
 * Candidates run the whole pipeline on a sample of the walks (every n/s-th
 * walk, keeping the mix of shapes): the batch itself up to max_walks walks,
 * halved until one run fits its share of budget_s. Each is timed by the wall
 * time of a plan run (best of two after a warm-up); candidates stop after
 * budget_s seconds (at least one is always measured); the default shape
 * (256 threads, 4096 steps) is kept unless a candidate beats it by more than
 * 3%. *best gets base with threads and chunk_steps set; *best_ms its
 * measured time. No counterpart in the reference: its CUDA
 * launches use a fixed BLOCK_SIZE of 256 (include/dreams/config.h:95,
 * src/cuda/persistent_kernel.cu:878-893). */
int drx_autotune(drx_ctx* ctx, const drx_batch* b, const drx_launch_cfg* base, int max_walks, double budget_s,
                 drx_launch_cfg* best, double* best_ms);

/* CRT of n rows of K canonical residues (row stride `stride`) over the
 * first K context primes: magnitude limbs (little-endian uint32, K per row)
 * and sign of centered(x, M). Rows of up to 256 primes take the wavefront
 * Garner kernel; wider rows (up to 8192 primes) the matrix-core CRT when at
 * least 64 rows of at most 1024 primes share the call, else the wide Garner
 * kernel (one workgroup per row). Replaces crt_reconstruct + centered
 * (runner.py:57-81). */
int drx_crt(drx_ctx* ctx, const int64_t* residues, int n, int K, int stride, uint32_t* limbs,
            uint8_t* neg);

/* drx_crt without any device allocation. No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same inputs, outputs and
 * results as drx_crt; every row takes a Garner kernel. Its device memory is
 * a fixed 8 MiB staging array per device, held until the process ends:
 * n * stride may be at most about 699000 residues (DRX_ERANGE above).
 * Calls that overlap on one device share that array and race; calls on one
 * context are serialised. Page-locked host buffers (drx_host_alloc) are
 * copied directly, pageable ones through the HIP runtime's staging. */
int drx_crt_noalloc(drx_ctx* ctx, const int64_t* residues, int n, int K, int stride, uint32_t* limbs,
                    uint8_t* neg);

/* Indices of the k largest values (ties: lower index first; NaN last).
 * No reference counterpart in Python; the reference CUDA sketch is
 * k_topk_select (src/cuda/persistent_kernel.cu:796-841). */
int drx_topk(drx_ctx* ctx, const double* values, int n, int k, int32_t* idx);

/* Page-locked host memory for batch outputs (and inputs): device-to-host
 * copies into it run at link speed with no staging. No counterpart in the
 * reference (its results are fresh numpy arrays). */
int drx_host_alloc(uint64_t bytes, void** out);
void drx_host_free(void* p);

/* Plain device copy of `bytes` (a multiple of 16) between two fresh buffers,
 * `reps` timed launches after one warm-up; *ms = mean device time per copy.
 * No counterpart in the reference: it is the calibration kernel for HBM byte
 * counters (each launch reads and writes exactly `bytes`) and the achievable
 * copy bandwidth the walk kernels are compared with. Allocates and frees. */
int drx_copy_calibrate(drx_ctx* ctx, uint64_t bytes, int reps, double* ms);

/* Issue rate of one VALU operation over the whole device: every lane runs
 * independent chains of op 0 (FP64 FMA), 1 (32-bit multiply-high), 2
 * (32x32+64-bit multiply-add) or 3 (the engine's residue modmul: FP64
 * two-product and quotient, counted as one operation). *ops_per_s gets the
 * measured rate, *peak_per_s CUs x 64 lanes x the device clock. No
 * counterpart in the reference: the measurement behind the FP64 residue
 * arithmetic of the walk kernels. */
int drx_valu_rate(drx_ctx* ctx, int op, double* ops_per_s, double* peak_per_s);

/* Device time of the last drx_walk_batch, from HIP events on the library's
 * stream: [0] walk kernels, [1] combine, [2] wait for the FP64 shadow chain
 * (which runs on a second stream beside the walk), [3] finalize + CRT +
 * top-k, [4] whole batch from the first walk kernel; milliseconds. [5]
 * walk-kernel launches. */
int drx_last_timing(const drx_ctx* ctx, double* t6);

/* ---- continuous-batching server ---------------------------------------
 * Requests are admitted into device-resident slots as memory allows; every
 * round advances each resident walk by `round_steps` steps or, while few
 * walks are resident, by up to `max_walk_tasks` such step ranges run side by
 * side (their ordered product is the same), so walks submitted while others
 * run join at the next round boundary. Finished walks get the GPU CRT of p
 * and q in the same round. No counterpart in the reference: its drivers call
 * run_pcf_walk + crt_reconstruct per request (python/euler2ai_full_verify.py:
 * 331-342, python/a100_sweep.py:416-436). */
typedef struct drx_server_cfg {
  int max_slots;        /* resident walks */
  int round_steps;      /* steps of one step range */
  int threads;          /* workgroup size, 0 = tuned */
  int max_walk_tasks;   /* step ranges per walk per round at light load (0 = 32) */
  int64_t mem_budget;   /* bytes of device state for slots (0 = 1 GiB) */
} drx_server_cfg;

typedef struct drx_server drx_server;

typedef struct drx_server_result {
  uint64_t ticket;
  drx_walk_out out;
  int32_t K;
  int32_t status;          /* 0 ok, < 0 error */
  double t_submit, t_admit, t_done;  /* seconds, steady clock; t_done: CRT integers in host memory */
  uint8_t neg[2];          /* sign of the centred CRT of p and q */
  uint8_t pad[6];
} drx_server_result;

int drx_server_create(drx_ctx* ctx, const drx_server_cfg* cfg, drx_server** out);
/* Copies the walk (and its axis values) into the server queue. */
int drx_server_submit(drx_server* s, const drx_walk* w, const int64_t* axis_vals, uint64_t* ticket);
/* Up to max finished results; residues[i] gets 2*K int64 (p then q) and,
 * when given, limbs[i] the CRT magnitudes of p and q (K little-endian uint32
 * each, as drx_crt), both [max][2][stride]; the signs are in res[i].neg. */
int drx_server_poll(drx_server* s, drx_server_result* res, int64_t* residues, uint32_t* limbs, int stride, int max,
                    int* n_done);
int drx_server_pending(drx_server* s, int* n_queued, int* n_resident);
/* Resident slots the memory budget (and max_slots) allow. */
int drx_server_capacity(drx_server* s, int* n_slots);
void drx_server_destroy(drx_server* s);

#ifdef __cplusplus
}
#endif
#endif /* DREAMS_HIP_H */
