// Walk kernels for gfx950: bytecode interpreter over residue channels, chunked
// P*M products with an ordered LDS tree combine, FP64 shadow, finalisation,
// Garner CRT and top-k.
#include "drx_kernels.h"

#include "drx_device.h"

namespace drx {

// ------------------------------------------------------------- interpreter
// One instruction stream per workgroup (the whole block walks one program),
// so fetch, decode and bank dispatch are wave-uniform scalar work. Banks are
// 16-wide vectors of doubles indexed with uniform indices (VGPR index mode).
typedef double v16d __attribute__((ext_vector_type(16)));
typedef double v4d __attribute__((ext_vector_type(4)));

// Entry bank: m*m doubles rounded up to a power of two.
template <int M> struct EBank { typedef v16d type; };
template <> struct EBank<2> { typedef v4d type; };

struct Lane {
  double p, pinv;
  uint32_t pe;
  uint32_t k;        // prime slot within the block's prime stripe
  uint32_t tid, T;   // for LDS temp slots
};

template <bool FP>
__device__ __forceinline__ double fetch(uint32_t bank, uint32_t i, const v16d& R, const v16d& X,
                                        const v16d& Cv, double acc, const double* __restrict__ ldc,
                                        uint32_t ldc_ld, const double* __restrict__ lds_s,
                                        const Lane& ln) {
  switch (bank) {
    case SB_R: return R[i & 15u];
    case SB_X: return X[i & 15u];
    case SB_C: return Cv[i & 15u];
    case SB_L: return FP ? ldc[i] : ldc[i * ldc_ld + ln.k];
    case SB_S: return lds_s[i * ln.T + ln.tid];
    case SB_A: return acc;
    default: return 0.0;
  }
}

template <bool FP, typename EV>
__device__ __forceinline__ void run_program(const uint2* __restrict__ code, uint32_t n, v16d& R,
                                            EV& E, const v16d& X, const v16d& Cv,
                                            const double* __restrict__ ldc, uint32_t ldc_ld,
                                            double* __restrict__ lds_s, const Lane& ln) {
  double acc = 0.0;
#pragma unroll 1
  for (uint32_t i = 0; i < n; ++i) {
    const uint2 w = code[i];
    const uint32_t op = w.x & 255u;
    const double x = fetch<FP>(w.y & 7u, (w.y >> 3) & 8191u, R, X, Cv, acc, ldc, ldc_ld, lds_s, ln);
    double r;
    if (op == D_MUL || op == D_ADD || op == D_SUB) {
      const double y = fetch<FP>((w.y >> 16) & 7u, w.y >> 19, R, X, Cv, acc, ldc, ldc_ld, lds_s, ln);
      if (op == D_MUL) r = FP ? x * y : mm(x, y, ln.p, ln.pinv);
      else r = (op == D_ADD) ? x + y : x - y;
    } else {
      switch (op) {
        case D_NEG: r = -x; break;
        case D_SQR: r = FP ? x * x : mm(x, x, ln.p, ln.pinv); break;
        case D_CUB: r = FP ? cube_cr(x) : mm(mm(x, x, ln.p, ln.pinv), x, ln.p, ln.pinv); break;
        case D_INV: r = FP ? finv(x) : inv_fermat(x, ln.p, ln.pinv, ln.pe); break;
        case D_RED: r = FP ? x : red(x, ln.p, ln.pinv); break;
        default: r = x; break;  // D_MOV
      }
    }
    const uint32_t db = (w.x >> 8) & 3u, d = w.x >> 10;
    if (db == DB_R) R[d & 15u] = r;
    else if (db == DB_E) E[d & (sizeof(EV) / sizeof(double) - 1)] = r;
    else if (db == DB_S) lds_s[d * ln.T + ln.tid] = r;
    acc = r;
  }
}

// P <- P * M where M's entries sit in E[0 .. M*M) and `kinds` marks the
// statically zero / one entries (uniform branches).
template <int M, typename EV>
__device__ __forceinline__ void step_matmul(double (&P)[M * M], const EV& E, uint32_t kinds,
                                            uint32_t colred, double p, double pinv) {
  double C[M * M];
#pragma unroll
  for (int i = 0; i < M; ++i) {
#pragma unroll
    for (int j = 0; j < M; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < M; ++k) {
        const uint32_t kd = (kinds >> (2 * (k * M + j))) & 3u;
        if (kd == E_GEN) acc += mm(P[i * M + k], E[k * M + j], p, pinv);
        else if (kd == E_ONE) acc += P[i * M + k];
      }
      if ((colred >> j) & 1u) acc = red(acc, p, pinv);
      C[i * M + j] = acc;
    }
  }
#pragma unroll
  for (int e = 0; e < M * M; ++e) P[e] = C[e];
}

// P <- P * Q, both general.
template <int M>
__device__ __forceinline__ void gen_matmul(double (&P)[M * M], const double (&Q)[M * M], double p,
                                           double pinv) {
  double C[M * M];
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < M; ++j) {
      double acc = mm(P[i * M], Q[j], p, pinv);
#pragma unroll
      for (int k = 1; k < M; ++k) acc += mm(P[i * M + k], Q[k * M + j], p, pinv);
      C[i * M + j] = acc;
    }
#pragma unroll
  for (int e = 0; e < M * M; ++e) P[e] = C[e];
}

__device__ __forceinline__ uint32_t pow2_at_least(uint32_t k) {
  uint32_t kp = 1;
  while (kp < k) kp <<= 1;
  return kp;
}

// ----------------------------------------------------------- walk kernels
// One workgroup per task (a step range of one walk). Lane (c, k): sub-chunk
// c of the range, prime k. Each lane multiplies M(s) over its sub-chunk; the
// C sub-chunk products are then combined in order by a log2(C)-level tree in
// LDS (the walk product is associative). Afterwards every lane evaluates the
// FP64 matrices M_f(s) for the shadow, strided over the whole range.
//
// LDS: [nconst][KP] constant residues + [nconst] FP64 constants (bank L),
// [nslots][T] temps (bank S), then the tree buffer [C][M*M][KP].

struct WalkSetup {
  DevTask t;
  DevWalk w;
  DevProgram pg;
  Lane ln;
  uint32_t K, KP, C, c, k;
  bool active;
  int64_t sb, se;
  const uint2* code;
  double *ldc, *ldcf, *lds_s, *tree;
};

template <int M>
__device__ __forceinline__ WalkSetup walk_setup(const KArgs& a, double* lds) {
  WalkSetup S;
  S.t = a.tasks[blockIdx.x];
  S.w = a.walks[S.t.walk];
  S.pg = a.progs[S.w.prog];
  const uint32_t T = blockDim.x, tid = threadIdx.x;
  S.K = (uint32_t)S.w.K;
  S.KP = pow2_at_least(S.K);
  S.C = T / S.KP;
  S.k = tid % S.KP;
  S.c = tid / S.KP;
  S.active = S.k < S.K;
  const uint32_t kk = S.active ? S.k : 0u;
  S.ln.p = a.prime_p[kk];
  S.ln.pinv = a.prime_pinv[kk];
  S.ln.pe = a.prime_e[kk];
  S.ln.k = S.k;
  S.ln.tid = tid;
  S.ln.T = T;
  S.code = reinterpret_cast<const uint2*>(a.code + S.pg.code_off);
  const uint32_t nl = S.pg.const_lds ? S.pg.nconst : 0u;
  S.ldc = lds;
  S.ldcf = S.ldc + (uint64_t)nl * S.KP;
  S.lds_s = S.ldcf + nl;
  S.tree = S.lds_s + (uint64_t)S.pg.nslots * T;
  for (uint32_t i = tid; i < nl * S.KP; i += T) {
    const uint32_t ci = i / S.KP, ck = i % S.KP;
    S.ldc[i] = (ck < S.K) ? a.cres[S.pg.cres_off + (uint64_t)ci * S.pg.kcap + ck] : 0.0;
  }
  for (uint32_t i = tid; i < nl; i += T) S.ldcf[i] = a.cflt[S.pg.cflt_off + i];
  __syncthreads();
  const int64_t L = S.t.s_end - S.t.s_begin;
  const int64_t len = (L + S.C - 1) / S.C;
  S.sb = S.t.s_begin + (int64_t)S.c * len;
  if (S.sb > S.t.s_end || !S.active) S.sb = S.t.s_end;
  S.se = S.sb + len;
  if (S.se > S.t.s_end) S.se = S.t.s_end;
  return S;
}

// Residue constants into the VGPR bank; axis values at step s0 and per-step
// increments into the X bank.
__device__ __forceinline__ void load_rns_inputs(const KArgs& a, const WalkSetup& S, int64_t s0, v16d& X,
                                                v16d& Cv) {
  const double p = S.ln.p, pinv = S.ln.pinv;
  const uint32_t kk = S.active ? S.k : 0u;
  if (!S.pg.const_lds) {
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if ((uint32_t)i < S.pg.nconst) Cv[i] = a.cres[S.pg.cres_off + (uint64_t)i * S.pg.kcap + kk];
  }
#pragma unroll
  for (int ax = 0; ax < kMaxDim; ++ax) {
    if ((uint32_t)ax < S.pg.dim) {
      const double* q = a.axr + S.w.axr_off + (uint64_t)ax * 3 * S.K;
      const double c0 = q[kk], c1 = q[S.K + kk], c2 = q[2 * S.K + kk];
      const double x0 = red(c0 + mm((double)s0, c1, p, pinv), p, pinv);
      X[ax] = mm(x0, c2, p, pinv);
      X[kIncSlot + ax] = mm(c1, c2, p, pinv);
    }
  }
}

__device__ __forceinline__ void advance_axes(v16d& X, uint32_t dim, double p, double pinv) {
#pragma unroll
  for (int ax = 0; ax < kMaxDim; ++ax)
    if ((uint32_t)ax < dim) X[ax] = red(X[ax] + X[kIncSlot + ax], p, pinv);
}

// Ordered tree combine of the C sub-chunk products; lane (0, k) ends with the
// product over the whole task and writes it to the chunk-product pool.
template <int M>
__device__ __forceinline__ void tree_combine(const KArgs& a, const WalkSetup& S, double (&P)[M * M]) {
  const uint32_t c = S.c, k = S.k, KP = S.KP, C = S.C;
  double* tree = S.tree;
#pragma unroll
  for (int e = 0; e < M * M; ++e) tree[((uint64_t)c * M * M + e) * KP + k] = P[e];
  for (uint32_t stride = 1; stride < C; stride <<= 1) {
    __syncthreads();
    if ((c % (2 * stride)) == 0 && c + stride < C) {
      double Q[M * M];
#pragma unroll
      for (int e = 0; e < M * M; ++e) Q[e] = tree[((uint64_t)(c + stride) * M * M + e) * KP + k];
      gen_matmul<M>(P, Q, S.ln.p, S.ln.pinv);
#pragma unroll
      for (int e = 0; e < M * M; ++e) tree[((uint64_t)c * M * M + e) * KP + k] = P[e];
    }
  }
  if (c == 0 && S.active) {
#pragma unroll
    for (int e = 0; e < M * M; ++e) a.cp[S.t.cp_off + (uint64_t)e * S.K + k] = P[e];
  }
}

// FP64 shadow matrices M_f(s) for every step of the task, one step per lane.
template <int M>
__device__ __forceinline__ void fp_pass(const KArgs& a, const WalkSetup& S) {
  const uint32_t T = blockDim.x, tid = threadIdx.x;
  typename EBank<M>::type E = 0.0;
  v16d R = 0.0, X = 0.0, Cv = 0.0;
  if (!S.pg.const_lds) {
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if ((uint32_t)i < S.pg.nconst) Cv[i] = a.cflt[S.pg.cflt_off + i];
  }
  for (int64_t s = S.t.s_begin + tid; s < S.t.s_end; s += T) {
#pragma unroll
    for (int ax = 0; ax < kMaxDim; ++ax) {
      if ((uint32_t)ax < S.pg.dim) {
        const int64_t* q = a.axi + S.w.axi_off + 2 * ax;
        X[ax] = (double)(q[0] + s * q[1]) / a.axd[S.w.axd_off + ax];
      }
    }
    X[kStepSlot] = (double)s;
    run_program<true>(S.code, S.pg.ncode, R, E, X, Cv, S.ldcf, 0, S.lds_s, S.ln);
    double* out = a.mf + S.t.mf_off + (uint64_t)(s - S.t.s_begin) * M * M;
#pragma unroll
    for (int e = 0; e < M * M; ++e) {
      const uint32_t kd = (S.pg.kinds >> (2 * e)) & 3u;
      out[e] = (kd == E_GEN) ? E[e] : (kd == E_ONE ? 1.0 : 0.0);
    }
  }
}

// Interpreting walk: the program runs once per step and residue channel.
template <int M>
__global__ void __launch_bounds__(256) k_walk_chunk(const KArgs a) {
  extern __shared__ double lds[];
  const WalkSetup S = walk_setup<M>(a, lds);
  const double p = S.ln.p, pinv = S.ln.pinv;
  v16d R = 0.0, X = 0.0, Cv = 0.0;
  typename EBank<M>::type E = 0.0;
  load_rns_inputs(a, S, S.sb, X, Cv);
  double P[M * M];
#pragma unroll
  for (int e = 0; e < M * M; ++e) P[e] = (e % (M + 1) == 0) ? 1.0 : 0.0;
  for (int64_t s = S.sb; s < S.se; ++s) {
    X[kStepSlot] = (double)s;
    run_program<false>(S.code, S.pg.ncode, R, E, X, Cv, S.ldc, S.KP, S.lds_s, S.ln);
    step_matmul<M>(P, E, S.pg.kinds, S.pg.colred, p, pinv);
    advance_axes(X, S.pg.dim, p, pinv);
  }
  tree_combine<M>(a, S, P);
  if (a.with_fp) fp_pass<M>(a, S);
}

// Forward-difference walk for programs whose general entries are all
// polynomials in the step (no INV on their paths). For an entry of degree
// d, the table D[e][0..d] holds its forward differences; one step is
// D[i] += D[i+1] for i < d (exact modular arithmetic, so the entry values are
// those the program would compute). The interpreter runs only to seed the
// tables: d + 1 evaluations at the start of each sub-chunk. Tables are
// reduced every fourth step, which keeps every entry below C(4 + d, d) * 2^31
// (< 2^41 for d <= DMAX).
template <int M, int DMAX>
__global__ void __launch_bounds__(256) k_walk_fd(const KArgs a) {
  extern __shared__ double lds[];
  const WalkSetup S = walk_setup<M>(a, lds);
  const double p = S.ln.p, pinv = S.ln.pinv;
  const uint32_t kinds = S.pg.kinds;
  const uint64_t degs = S.pg.fd_deg;
  double D[M * M][DMAX + 1];
  {
    v16d R = 0.0, X = 0.0, Cv = 0.0;
    typename EBank<M>::type E = 0.0;
    load_rns_inputs(a, S, S.sb, X, Cv);
#pragma unroll
    for (int j = 0; j <= DMAX; ++j) {
      if ((uint32_t)j <= S.pg.fd_degree) {
        X[kStepSlot] = (double)(S.sb + j);
        run_program<false>(S.code, S.pg.ncode, R, E, X, Cv, S.ldc, S.KP, S.lds_s, S.ln);
        advance_axes(X, S.pg.dim, p, pinv);
      }
#pragma unroll
      for (int e = 0; e < M * M; ++e) D[e][j] = ((uint32_t)j <= S.pg.fd_degree) ? E[e] : 0.0;
    }
  }
  // Newton forward differences: D[e][j] <- Delta^j f(sb)
#pragma unroll
  for (int i = 1; i <= DMAX; ++i)
#pragma unroll
    for (int j = DMAX; j >= i; --j)
#pragma unroll
      for (int e = 0; e < M * M; ++e) D[e][j] = D[e][j] - D[e][j - 1];
#pragma unroll
  for (int e = 0; e < M * M; ++e)
#pragma unroll
    for (int j = 0; j <= DMAX; ++j) D[e][j] = red(D[e][j], p, pinv);

  double P[M * M];
#pragma unroll
  for (int e = 0; e < M * M; ++e) P[e] = (e % (M + 1) == 0) ? 1.0 : 0.0;
  int64_t s = S.sb;
  uint32_t phase = 0;
  for (; s < S.se; ++s) {
    double Cm[M * M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
#pragma unroll
      for (int j = 0; j < M; ++j) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < M; ++q) {
          const uint32_t kd = (kinds >> (2 * (q * M + j))) & 3u;
          if (kd == E_GEN) acc += mm(P[i * M + q], D[q * M + j][0], p, pinv);
          else if (kd == E_ONE) acc += P[i * M + q];
        }
        if ((S.pg.colred >> j) & 1u) acc = red(acc, p, pinv);
        Cm[i * M + j] = acc;
      }
    }
#pragma unroll
    for (int e = 0; e < M * M; ++e) P[e] = Cm[e];
    // advance the tables of the general entries
    phase = (phase + 1) & 3u;
#pragma unroll
    for (int e = 0; e < M * M; ++e) {
      if (((kinds >> (2 * e)) & 3u) != E_GEN) continue;
      const uint32_t d = (uint32_t)(degs >> (4 * e)) & 15u;
#pragma unroll
      for (int i = 0; i < DMAX; ++i)
        if ((uint32_t)i < d) D[e][i] += D[e][i + 1];
      if (phase == 0) {
#pragma unroll
        for (int i = 0; i < DMAX; ++i)
          if ((uint32_t)i < d) D[e][i] = red(D[e][i], p, pinv);
      }
    }
  }
  tree_combine<M>(a, S, P);
  if (a.with_fp) fp_pass<M>(a, S);
}

// --------------------------------------------------------- ordered combine
// state <- state * cp[t0] * cp[t1] * ... over the walk's tasks (step order).
template <int M>
__global__ void k_combine(const KArgs a) {
  const DevSpan sp = a.spans[blockIdx.x];
  const DevWalk w = a.walks[sp.walk];
  const uint32_t K = (uint32_t)w.K;
  for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) {
    const double p = a.prime_p[k], pinv = a.prime_pinv[k];
    double S[M * M];
#pragma unroll
    for (int e = 0; e < M * M; ++e) S[e] = a.res_in[w.res_off + (uint64_t)e * K + k];
    for (int ti = sp.task_begin; ti < sp.task_end; ++ti) {
      const uint64_t off = a.tasks[ti].cp_off;
      double Q[M * M];
#pragma unroll
      for (int e = 0; e < M * M; ++e) Q[e] = a.cp[off + (uint64_t)e * K + k];
      gen_matmul<M>(S, Q, p, pinv);
    }
#pragma unroll
    for (int e = 0; e < M * M; ++e) a.res[w.res_off + (uint64_t)e * K + k] = S[e];
  }
}

// ------------------------------------------------------------ FP64 shadow
// One lane per walk: P_f <- P_f @ M_f(s) for every step in order, each entry
// a sequential FMA chain acc = a0*b0; acc = fma(ak, bk, acc); then
// mx = max|P_f| (NaN propagates) and, if mx > 1e10, P_f /= mx and
// log_scale += log(mx). Finishes with delta_float against the walk target.
template <int M>
__global__ void k_shadow(const KArgs a, int nspan, int finalize) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nspan) return;
  const DevSpan sp = a.spans[i];
  const DevWalk w = a.walks[sp.walk];
  const double* st_in = a.fst_in + w.fst_off;
  double* st = a.fst + w.fst_off;
  double P[M * M];
#pragma unroll
  for (int e = 0; e < M * M; ++e) P[e] = st_in[e];
  double ls = st_in[M * M];
  // The loop body is one branch-free block: the rescale is a select, and
  // log(mx) of step s is added while step s + 1 is multiplied (it only feeds
  // log_scale, in the same order as the reference adds it).
  double pend = 1.0;  // mx awaiting its log; log(1) == 0 exactly
  for (int ti = sp.task_begin; ti < sp.task_end; ++ti) {
    const DevTask t = a.tasks[ti];
    const double* mf = a.mf + t.mf_off;
    for (int64_t s = t.s_begin; s < t.s_end; ++s, mf += M * M) {
      double Mf[M * M];
#pragma unroll
      for (int e = 0; e < M * M; ++e) Mf[e] = mf[e];
      double C[M * M];
#pragma unroll
      for (int r = 0; r < M; ++r)
#pragma unroll
        for (int cc = 0; cc < M; ++cc) {
          double acc = P[r * M] * Mf[cc];
#pragma unroll
          for (int q = 1; q < M; ++q) acc = __builtin_fma(P[r * M + q], Mf[q * M + cc], acc);
          C[r * M + cc] = acc;
        }
      ls += log(pend);
      double mx = 0.0;
      bool nan = false;
#pragma unroll
      for (int e = 0; e < M * M; ++e) {
        const double v = __builtin_fabs(C[e]);
        nan |= (v != v);
        mx = (v > mx) ? v : mx;
      }
      const bool resc = !nan && mx > 1e10;
      pend = resc ? mx : 1.0;
#pragma unroll
      for (int e = 0; e < M * M; ++e) P[e] = resc ? C[e] / mx : C[e];
    }
  }
  ls += log(pend);
#pragma unroll
  for (int e = 0; e < M * M; ++e) st[e] = P[e];
  st[M * M] = ls;
  if (finalize) {
    // delta_float (runner.py:86-100)
    const double pf = P[M - 1], qf = P[M * M - 1];
    double delta, logq;
    if (!isfinite(pf) || !isfinite(qf) || __builtin_fabs(qf) < 1e-300) {
      delta = -1e10;
      logq = 0.0;
    } else {
      const double err = __builtin_fabs(pf / qf - w.target);
      logq = ls + log(__builtin_fabs(qf));
      if (err < 1e-300) delta = 100.0;
      else if (logq <= 0.0) delta = -1e10;
      else delta = -(1.0 + log(err) / logq);
    }
    double* o = a.fout + (uint64_t)sp.walk * 6;
    o[0] = pf;
    o[1] = qf;
    o[2] = ls;
    o[3] = delta;
    o[4] = logq;
  }
}

// --------------------------------------------------------------- finalize
// Canonical residues of the last column (p = P[0][m-1], q = P[m-1][m-1])
// and, when requested, of the whole state matrix.
__global__ void k_finalize(const KArgs a, int full) {
  const DevSpan sp = a.spans[blockIdx.x];
  const DevWalk w = a.walks[sp.walk];
  const uint32_t K = (uint32_t)w.K, m = (uint32_t)w.m;
  const double* S = a.res + w.res_off;
  int64_t* ro = a.rout + (uint64_t)sp.out_row * 2 * a.out_ld;
  if (a.fcompact && threadIdx.x < 6)
    a.fcompact[(uint64_t)sp.out_row * 6 + threadIdx.x] = a.fout[(uint64_t)sp.walk * 6 + threadIdx.x];
  if (a.ffull && threadIdx.x < 16)
    a.ffull[(uint64_t)sp.out_row * 16 + threadIdx.x] = threadIdx.x < m * m ? a.fst[w.fst_off + threadIdx.x] : 0.0;
  for (uint32_t k = threadIdx.x; k < K; k += blockDim.x) {
    const double p = a.prime_p[k], pinv = a.prime_pinv[k];
    ro[k] = (int64_t)canon(S[(uint64_t)(m - 1) * K + k], p, pinv);
    ro[a.out_ld + k] = (int64_t)canon(S[(uint64_t)(m * m - 1) * K + k], p, pinv);
    if (full) {
      int64_t* po = a.pout + (uint64_t)sp.out_row * 16 * a.out_ld;
      for (uint32_t e = 0; e < m * m; ++e)
        po[(uint64_t)e * a.out_ld + k] = (int64_t)canon(S[(uint64_t)e * K + k], p, pinv);
    }
  }
}

// -------------------------------------------------------------- Garner CRT
// One thread per (walk, value): mixed-radix digits v_i of x in [0, M), the
// sign of the centred value (x > (M-1)/2 <=> digits compare above
// (p_i - 1)/2 from the top), the digits of M - x when negative, and the
// binary limbs by Horner x = (...(v_{K-1} p_{K-2} + v_{K-2}) ...) p_0 + v_0.
// One wavefront per row. Digit / limb t of the row lives in lane t % 64,
// register t / 64 (J registers, K <= 64 J). Mixed-radix digits by the
// recurrence t_i <- (t_i - v_j) p_j^{-1} mod p_i, v_j broadcast from its lane;
// then the sign against (p_i - 1) / 2 from the top digit, the negation
// M - x for the upper half, and Horner x = v_0 + p_0 (v_1 + p_1 (...)) on
// 32-bit limbs with a ballot carry-lookahead across lanes.
template <int J>
__global__ void __launch_bounds__(256) k_garner_wave(const GarnerArgs g) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= g.n) return;  // the whole wavefront leaves together
  const int K = g.row_k ? g.row_k[row] : g.K;
  const int64_t* r = g.res + (uint64_t)row * g.res_stride;
  double t[J], pd[J], pv[J];
  uint32_t pu[J], v[J];
#pragma unroll
  for (int jj = 0; jj < J; ++jj) {
    const int i = jj * 64 + lane;
    const bool in = i < K;
    pd[jj] = in ? g.prime_p[i] : 3.0;
    pv[jj] = in ? g.prime_pinv[i] : 1.0 / 3.0;
    pu[jj] = in ? g.prime_u[i] : 3u;
    t[jj] = in ? (double)r[i] : 0.0;
  }
  // mixed-radix digits: after step j, t_j is final
#pragma unroll
  for (int jr = 0; jr < J; ++jr) {
    const int jbase = jr * 64;
    const int lim = (K - jbase < 64) ? K - jbase : 64;
    for (int jl = 0; jl < lim; ++jl) {
      const int j = jbase + jl;
      const double vj = (double)__builtin_amdgcn_readlane((int)t[jr], jl);
      const uint32_t* inv = g.inv_table + (uint64_t)j * g.inv_ld;
#pragma unroll
      for (int jj = jr; jj < J; ++jj) {
        const int i = jj * 64 + lane;
        const bool upd = i > j && i < K;
        const double w = upd ? (double)inv[i] : 0.0;
        const double x = canon(mm(t[jj] - vj, w, pd[jj], pv[jj]), pd[jj], pv[jj]);
        t[jj] = upd ? x : t[jj];
      }
    }
  }
#pragma unroll
  for (int jj = 0; jj < J; ++jj) v[jj] = (jj * 64 + lane < K) ? (uint32_t)t[jj] : 0u;

  // sign: the most significant digit that differs from (p_i - 1) / 2
  int neg = 0, found = 0;
#pragma unroll
  for (int jj = J - 1; jj >= 0; --jj) {
    const uint32_t h = (pu[jj] - 1u) >> 1;
    const uint64_t m = __ballot((jj * 64 + lane < K) && v[jj] != h);
    if (!found && m) {
      const int top = 63 - __clzll((long long)m);
      neg = __builtin_amdgcn_readlane(v[jj] > h ? 1 : 0, top);
      found = 1;
    }
  }
  if (neg) {  // digits of M - x = (M - 1 - x) + 1: the +1 ripples through zero digits
    int below_nz = 0;
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
      const bool in = jj * 64 + lane < K;
      const uint64_t nz = __ballot(in && v[jj] != 0u);
      const bool carry = !below_nz && (nz & ((1ull << lane) - 1ull)) == 0ull;
      const uint32_t d = carry ? (v[jj] == 0u ? 0u : pu[jj] - v[jj]) : pu[jj] - 1u - v[jj];
      v[jj] = in ? d : 0u;
      below_nz = below_nz || nz != 0ull;
    }
  }

  // Horner from the top digit into little-endian 32-bit limbs
  uint32_t L[J];
#pragma unroll
  for (int jj = 0; jj < J; ++jj) L[jj] = 0u;
  {
    const int top = K - 1;
    uint32_t vt = 0u;
#pragma unroll
    for (int jj = 0; jj < J; ++jj)
      if (jj == top / 64) vt = (uint32_t)__builtin_amdgcn_readlane((int)v[jj], top % 64);
    L[0] = (lane == 0) ? vt : 0u;
  }
#pragma unroll
  for (int jr = J - 1; jr >= 0; --jr) {
    const int ib = jr * 64;
    const int hi_l = (K - 2 - ib < 63) ? K - 2 - ib : 63;
    for (int jl = hi_l; jl >= 0; --jl) {
      const uint32_t p = (uint32_t)__builtin_amdgcn_readlane((int)pu[jr], jl);
      const uint32_t vi = (uint32_t)__builtin_amdgcn_readlane((int)v[jr], jl);
      uint32_t lo[J], hi[J], c[J];
#pragma unroll
      for (int jj = 0; jj < J; ++jj) {
        const uint64_t z = (uint64_t)L[jj] * p;
        lo[jj] = (uint32_t)z;
        hi[jj] = (uint32_t)(z >> 32);
      }
      // limb t gets lo_t + hi_{t-1} (+ v_i at t = 0): below 2^33
#pragma unroll
      for (int jj = 0; jj < J; ++jj) {
        uint32_t hp = (uint32_t)__shfl_up((int)hi[jj], 1);
        const uint32_t h0 = (jj == 0) ? vi : (uint32_t)__builtin_amdgcn_readlane((int)hi[jj > 0 ? jj - 1 : 0], 63);
        hp = (lane == 0) ? h0 : hp;
        const uint64_t X = (uint64_t)lo[jj] + hp;
        L[jj] = (uint32_t)X;
        c[jj] = (uint32_t)(X >> 32);
      }
      // then + c_{t-1}, with the 1-bit ripple resolved by carry-lookahead
      uint64_t Y[J], gm[J], pm[J];
#pragma unroll
      for (int jj = 0; jj < J; ++jj) {
        uint32_t cp = (uint32_t)__shfl_up((int)c[jj], 1);
        const uint32_t c0 = (jj == 0) ? 0u : (uint32_t)__builtin_amdgcn_readlane((int)c[jj > 0 ? jj - 1 : 0], 63);
        cp = (lane == 0) ? c0 : cp;
        Y[jj] = (uint64_t)L[jj] + cp;
        gm[jj] = __ballot((Y[jj] >> 32) != 0ull);
        pm[jj] = __ballot(Y[jj] == 0xFFFFFFFFull);
      }
      uint64_t cin = 0ull;
#pragma unroll
      for (int jj = 0; jj < J; ++jj) {
        const uint64_t A = gm[jj] | pm[jj], B = gm[jj];
        const uint64_t S1 = A + B;
        const uint64_t S = S1 + cin;
        const uint64_t CI = S ^ A ^ B;  // carry into each lane's limb
        cin = (S1 < A || S < S1) ? 1ull : 0ull;
        L[jj] = (uint32_t)(Y[jj] + ((CI >> lane) & 1ull));
      }
    }
  }
  uint32_t* limb = g.limbs + (uint64_t)row * g.res_stride;
#pragma unroll
  for (int jj = 0; jj < J; ++jj)
    if (jj * 64 + lane < K) limb[jj * 64 + lane] = L[jj];
  if (lane == 0) g.neg[row] = (uint8_t)neg;
}

// ------------------------------------------------------------------ top-k
// Order: larger delta first, NaN last, ties to the lower index. A double
// maps to a 64-bit key whose unsigned order is the numeric order (-0 and +0
// equal, NaN below everything); (key desc, index asc) is then a strict total
// order on the walks.
__device__ __forceinline__ uint64_t order_key(double d) {
  if (d != d) return 0ull;
  const uint64_t b = (d == 0.0) ? 0ull : (uint64_t)__double_as_longlong(d);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ bool better(double dj, int j, double di, int i) {
  const bool nj = dj != dj, ni = di != di;
  if (nj != ni) return ni;            // non-NaN beats NaN
  if (!nj && dj != di) return dj > di;
  return j < i;
}

// Tournament round: each block sorts kTopB candidates in LDS (bitonic, in
// the order above) and keeps its first kk. Padding candidates carry key 0 and
// index 0xFFFFFFFF, after every real one. The last round (one block) writes
// the winners' indices.
constexpr int kTopB = 2048;

template <bool FROM_VALUES>
__global__ void __launch_bounds__(256) k_topk_round(const double* __restrict__ d, int stride,
                                                    const TopCand* __restrict__ cin, int n, int kk,
                                                    TopCand* __restrict__ cout, int32_t* __restrict__ out) {
  __shared__ uint64_t sk[kTopB];
  __shared__ uint32_t si[kTopB];
  const int base = blockIdx.x * kTopB;
  for (int t = threadIdx.x; t < kTopB; t += 256) {
    const int i = base + t;
    uint64_t key = 0ull;
    uint32_t idx = 0xFFFFFFFFu;
    if (i < n) {
      if (FROM_VALUES) {
        key = order_key(d[(uint64_t)i * stride]);
        idx = (uint32_t)i;
      } else {
        key = cin[i].key;
        idx = cin[i].idx;
      }
    }
    sk[t] = key;
    si[t] = idx;
  }
  __syncthreads();
  for (int size = 2; size <= kTopB; size <<= 1) {
    for (int half = size >> 1; half > 0; half >>= 1) {
      for (int t = threadIdx.x; t < kTopB / 2; t += 256) {
        const int lo = 2 * t - (t & (half - 1));
        const int hi = lo + half;
        const uint64_t ka = sk[lo], kb = sk[hi];
        const uint32_t ia = si[lo], ib = si[hi];
        const bool b_first = kb > ka || (kb == ka && ib < ia);
        if (b_first == ((lo & size) == 0)) {
          sk[lo] = kb;
          sk[hi] = ka;
          si[lo] = ib;
          si[hi] = ia;
        }
      }
      __syncthreads();
    }
  }
  if (gridDim.x == 1 && out) {
    for (int t = threadIdx.x; t < kk; t += 256) out[t] = (int32_t)si[t];
  } else {
    for (int t = threadIdx.x; t < kk; t += 256) {
      TopCand c;
      c.key = sk[t];
      c.idx = si[t];
      c.pad = 0u;
      cout[(uint64_t)blockIdx.x * kk + t] = c;
    }
  }
}

// k above the tournament's reach: rank(i) = #{j before i}, element i written
// at out[rank] when rank < k. LDS-tiled O(n^2 / T).
__global__ void k_topk_rank(const double* __restrict__ d, int stride, int n, int k,
                            int32_t* __restrict__ out) {
  __shared__ double tile[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double di = (i < n) ? d[(uint64_t)i * stride] : 0.0;
  int rank = 0;
  for (int base = 0; base < n; base += 256) {
    __syncthreads();
    if (base + (int)threadIdx.x < n) tile[threadIdx.x] = d[(uint64_t)(base + threadIdx.x) * stride];
    __syncthreads();
    const int lim = (n - base < 256) ? n - base : 256;
    for (int j = 0; j < lim; ++j) rank += better(tile[j], base + j, di, i) ? 1 : 0;
  }
  if (i < n && rank < k) out[rank] = i;
}

// ---------------------------------------------------------------- launchers
hipError_t launch_walk(int m, int fd, const KArgs& a, int ntask, int threads, size_t lds, hipStream_t s) {
  if (ntask <= 0) return hipSuccess;
  if (fd) {
    if (m == 2) hipLaunchKernelGGL((k_walk_fd<2, kFdMax2>), dim3(ntask), dim3(threads), lds, s, a);
    else if (m == 3) hipLaunchKernelGGL((k_walk_fd<3, kFdMax3>), dim3(ntask), dim3(threads), lds, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
  }
  if (m == 2) hipLaunchKernelGGL(k_walk_chunk<2>, dim3(ntask), dim3(threads), lds, s, a);
  else if (m == 3) hipLaunchKernelGGL(k_walk_chunk<3>, dim3(ntask), dim3(threads), lds, s, a);
  else if (m == 4) hipLaunchKernelGGL(k_walk_chunk<4>, dim3(ntask), dim3(threads), lds, s, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_combine(int m, const KArgs& a, int nspan, hipStream_t s) {
  if (nspan <= 0) return hipSuccess;
  if (m == 2) hipLaunchKernelGGL(k_combine<2>, dim3(nspan), dim3(64), 0, s, a);
  else if (m == 3) hipLaunchKernelGGL(k_combine<3>, dim3(nspan), dim3(64), 0, s, a);
  else if (m == 4) hipLaunchKernelGGL(k_combine<4>, dim3(nspan), dim3(64), 0, s, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_shadow(int m, const KArgs& a, int nspan, int finalize, hipStream_t s) {
  if (nspan <= 0) return hipSuccess;
  const int tb = 64, nb = (nspan + tb - 1) / tb;
  if (m == 2) hipLaunchKernelGGL(k_shadow<2>, dim3(nb), dim3(tb), 0, s, a, nspan, finalize);
  else if (m == 3) hipLaunchKernelGGL(k_shadow<3>, dim3(nb), dim3(tb), 0, s, a, nspan, finalize);
  else if (m == 4) hipLaunchKernelGGL(k_shadow<4>, dim3(nb), dim3(tb), 0, s, a, nspan, finalize);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_finalize(const KArgs& a, int nspan, int full, hipStream_t s) {
  if (nspan <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize, dim3(nspan), dim3(64), 0, s, a, full);
  return hipGetLastError();
}

hipError_t launch_garner(const GarnerArgs& g, hipStream_t s) {
  if (g.n <= 0) return hipSuccess;
  const dim3 grid((g.n + 3) / 4), block(256);
  if (g.K <= 64) hipLaunchKernelGGL(k_garner_wave<1>, grid, block, 0, s, g);
  else if (g.K <= 128) hipLaunchKernelGGL(k_garner_wave<2>, grid, block, 0, s, g);
  else if (g.K <= 256) hipLaunchKernelGGL(k_garner_wave<4>, grid, block, 0, s, g);
  else if (g.K <= kMaxK) hipLaunchKernelGGL(k_garner_wave<kMaxK / 64>, grid, block, 0, s, g);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

// Plain 16-byte-vector copy: the calibration kernel for HBM byte counters
// (it moves exactly 2 x n16 x 16 bytes) and the achievable-bandwidth probe.
__global__ void __launch_bounds__(256) k_copy_calib(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                    size_t n16) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

hipError_t launch_copy(const void* src, void* dst, size_t n16, hipStream_t s) {
  if (n16 == 0) return hipSuccess;
  const size_t tb = 256, want = (n16 + tb - 1) / tb;
  const unsigned nb = (unsigned)(want < 8192 ? want : 8192);
  hipLaunchKernelGGL(k_copy_calib, dim3(nb), dim3(tb), 0, s, (const uint4*)src, (uint4*)dst, n16);
  return hipGetLastError();
}

static int topk_keep(int k) { return k < kTopB / 2 ? k : kTopB / 2; }

size_t topk_workspace_bytes(int n, int k) {
  if (n <= kTopB || k > kTopB / 2) return 0;
  const size_t nb = ((size_t)n + kTopB - 1) / kTopB;
  return 2 * nb * (size_t)topk_keep(k) * sizeof(TopCand);
}

hipError_t launch_topk(const double* d, int stride, int n, int k, int32_t* out, void* ws, hipStream_t s) {
  if (n <= 0 || k <= 0) return hipSuccess;
  if (k > n) return hipErrorInvalidValue;
  if (k > kTopB / 2) {
    hipLaunchKernelGGL(k_topk_rank, dim3((n + 255) / 256), dim3(256), 0, s, d, stride, n, k, out);
    return hipGetLastError();
  }
  const int kk = topk_keep(k);
  int cur = n;
  int nb = (cur + kTopB - 1) / kTopB;
  if (nb == 1) {
    hipLaunchKernelGGL(k_topk_round<true>, dim3(1), dim3(256), 0, s, d, stride, nullptr, cur, k, nullptr, out);
    return hipGetLastError();
  }
  if (!ws) return hipErrorInvalidValue;
  TopCand* buf[2] = {(TopCand*)ws, (TopCand*)ws + (size_t)nb * kk};
  hipLaunchKernelGGL(k_topk_round<true>, dim3(nb), dim3(256), 0, s, d, stride, nullptr, cur, kk, buf[0], nullptr);
  cur = nb * kk;
  int src = 0;
  for (;;) {
    nb = (cur + kTopB - 1) / kTopB;
    if (nb == 1) {
      hipLaunchKernelGGL(k_topk_round<false>, dim3(1), dim3(256), 0, s, nullptr, 0, buf[src], cur, k, nullptr, out);
      break;
    }
    hipLaunchKernelGGL(k_topk_round<false>, dim3(nb), dim3(256), 0, s, nullptr, 0, buf[src], cur, kk, buf[1 - src],
                       nullptr);
    cur = nb * kk;
    src = 1 - src;
  }
  return hipGetLastError();
}

}  // namespace drx
