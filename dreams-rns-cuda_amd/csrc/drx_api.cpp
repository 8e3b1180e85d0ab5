// C ABI of libdreams_hip: context, primes, programs, batch plans, CRT, top-k.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <numeric>

#include "drx_crt.h"
#include "drx_ctx.h"
#include "drx_device.h"

using namespace drx;

namespace drx {

int fail(drx_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

#define HIPCHK DRX_HIPCHK

// ---- host modular arithmetic (p < 2^31)

bool is_prime_u32(uint32_t n) {
  if (n < 2) return false;
  for (uint32_t q : {2u, 3u, 5u, 7u, 11u, 13u}) {
    if (n % q == 0) return n == q;
  }
  uint32_t d = n - 1;
  int s = 0;
  while ((d & 1) == 0) {
    d >>= 1;
    ++s;
  }
  for (uint32_t a : {2u, 3u, 5u, 7u}) {  // deterministic below 3.2e9
    uint64_t x = powmod(a, d, n);
    if (x == 1 || x == n - 1) continue;
    bool comp = true;
    for (int r = 1; r < s; ++r) {
      x = x * x % n;
      if (x == n - 1) {
        comp = false;
        break;
      }
    }
    if (comp) return false;
  }
  return true;
}


double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int upload_programs(drx_ctx* ctx) {
  if (!ctx->progs_dirty) return DRX_OK;
  const size_t np = ctx->progs.size();
  std::vector<DevProgram> dp(np);
  for (size_t i = 0; i < np; ++i) dp[i] = ctx->progs[i].dev;
  // padded past the end by a margin
  const size_t code_words = ctx->h_code.size() + 2 * kCodeFetch;
  HIPCHK(ctx, ctx->d_code.ensure_keep(code_words * 4, ctx->retired));
  HIPCHK(ctx, ctx->d_cres.ensure_keep(std::max<size_t>(1, ctx->h_cres.size()) * 8, ctx->retired));
  HIPCHK(ctx, ctx->d_cflt.ensure_keep(std::max<size_t>(1, ctx->h_cflt.size()) * 8, ctx->retired));
  HIPCHK(ctx, ctx->d_progs.ensure_keep(std::max<size_t>(1, np) * sizeof(DevProgram), ctx->retired));
  {
    std::vector<uint32_t> padded(code_words, 0u);
    std::copy(ctx->h_code.begin(), ctx->h_code.end(), padded.begin());
    HIPCHK(ctx, hipMemcpy(ctx->d_code.p, padded.data(), code_words * 4, hipMemcpyHostToDevice));
  }
  if (!ctx->h_cres.empty())
    HIPCHK(ctx, hipMemcpy(ctx->d_cres.p, ctx->h_cres.data(), ctx->h_cres.size() * 8, hipMemcpyHostToDevice));
  if (!ctx->h_cflt.empty())
    HIPCHK(ctx, hipMemcpy(ctx->d_cflt.p, ctx->h_cflt.data(), ctx->h_cflt.size() * 8, hipMemcpyHostToDevice));
  if (np) HIPCHK(ctx, hipMemcpy(ctx->d_progs.p, dp.data(), np * sizeof(DevProgram), hipMemcpyHostToDevice));
  ctx->progs_dirty = false;
  return DRX_OK;
}

// p_j^{-1} mod p_i for the first k primes (lazily extended).
int walk_kind(const HostProgram& hp, bool interp_only) {
  if (interp_only || !hp.fd) return WALK_INTERP;
  if (pcf_shape(hp) >= 0) return WALK_FD_PCF;
  return comp_shape(hp.dev) >= 0 ? WALK_FD_COMP : WALK_FD;
}

int walk_shape(const HostProgram& hp, int kind) {
  if (kind == WALK_FD_PCF) return pcf_shape(hp);
  if (kind == WALK_FD_COMP) return comp_shape(hp.dev);
  return 0;
}

size_t device_lds_limit(drx_ctx* ctx) {
  // less 64 bytes: the persistent walk kernels' task word is static LDS beside the dynamic buffer
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) == hipSuccess && v > 64)
    return (size_t)v - 64;
  return 65536 - 64;
}

void merge_lds_classes(std::vector<Stage>& stages) {
  // adjacent stages of one walk kernel run as one when the larger LDS keeps
  // the smaller one's occupancy: a separate launch would only add a tail
  std::vector<Stage> out;
  for (const Stage& g : stages) {
    if (!out.empty()) {
      Stage& a = out.back();
      if (a.m == g.m && a.kind == g.kind && a.kind == WALK_INTERP && a.threads == g.threads &&
          a.task_end == g.task_begin && a.span_end == g.span_begin && a.range_end == g.range_begin &&
          a.seed_end == g.seed_begin) {
        const size_t lmax = std::max(a.lds, g.lds);
        const int oa = walk_blocks_per_cu(a.m, a.kind, a.threads, std::min(a.lds, g.lds));
        const int om = walk_blocks_per_cu(a.m, a.kind, a.threads, lmax);
        if (oa > 0 && om == oa) {
          a.task_end = g.task_end;
          a.span_end = g.span_end;
          a.range_end = g.range_end;
          a.seed_end = g.seed_end;
          a.lds = lmax;
          a.mf_lds = std::max(a.mf_lds, g.mf_lds);
          a.seed_lds = std::max(a.seed_lds, g.seed_lds);
          a.seed_kp = std::max(a.seed_kp, g.seed_kp);
          continue;
        }
      }
    }
    out.push_back(g);
  }
  stages.swap(out);
}

int lds_class(const HostProgram& hp, int kind) {
  return kind == WALK_INTERP ? std::min<int>(3, ((int)hp.dev.nslots + 7) / 8) : 0;
}

int pcf_shape(const HostProgram& hp) {
  // companion PCF [[0, b], [1, a]]: entry (0, 0) zero, (1, 0) one; b and a of
  // any kind (a static 0 or 1 is a table of degree 0)
  if (!hp.fd || hp.dev.m != 2) return -1;
  const uint32_t k = hp.dev.kinds;
  if ((k & 3u) != E_ZERO || ((k >> 4) & 3u) != E_ONE) return -1;
  const int db = ((k >> 2) & 3u) == E_GEN ? (int)((hp.dev.fd_deg >> 4) & 15u) : 0;
  const int da = ((k >> 6) & 3u) == E_GEN ? (int)((hp.dev.fd_deg >> 12) & 15u) : 0;
  if (db > kFdMax2 || da > kFdMax2) return -1;
  return 8 * db + da;
}

int comp_shape(const DevProgram& d) {
  // companion CMF: ones on the sub-diagonal, zeros elsewhere left of the last
  // column; the last column polynomial (a static 0 or 1 is degree 0)
  const int m = (int)d.m;
  if (!d.fd_ok || m < 3 || m > 4) return -1;
  int deg = 0;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      const int e = i * m + j;
      const uint32_t kd = (d.kinds >> (2 * e)) & 3u;
      if (j < m - 1) {
        if (kd != (i == j + 1 ? (uint32_t)E_ONE : (uint32_t)E_ZERO)) return -1;
      } else if (kd == E_GEN) {
        deg = std::max(deg, (int)((d.fd_deg >> (4 * e)) & 15u));
      }
    }
  return deg <= kFdMax3 ? deg : -1;
}

// Constant slots of the residue code beyond those held in the VGPR bank C.
static size_t lds_const_rows(const DevProgram& pg) {
  const uint32_t nc = (uint32_t)kResStep - pg.dim;
  return pg.nconst > nc ? pg.nconst - nc : 0;
}

size_t walk_lds_bytes(const DevProgram& pg, int kind, int m, uint32_t K, int threads, int wpg, int slice) {
  const size_t tree = (size_t)threads * m * m;  // [W][C][m*m][KP] with W * C * KP = threads
  if (kind != WALK_INTERP) return 8 * tree;
  const size_t KP = pow2_ceil(std::min<uint32_t>(K, (uint32_t)slice));  // one prime slice
  const size_t nl = lds_const_rows(pg);
  // constant rows by prime, axis-increment rows by lane group and prime,
  // spill and entry rows by thread; the tree overlays them
  return 8 * std::max((nl + (size_t)std::max(1, wpg) * pg.dim) * KP +
                          ((size_t)pg.nslots + (size_t)m * m) * threads, tree);
}

size_t seed_lds_bytes(const DevProgram& pg, int kp, int seed_threads) {
  const size_t nl = lds_const_rows(pg);
  return 8 * ((nl + pg.dim) * (size_t)kp + ((size_t)pg.nslots + (size_t)pg.m * pg.m) * seed_threads);
}

size_t mf_lds_bytes(const DevProgram& pg) {
  return 8 * ((size_t)pg.nconst + (size_t)pg.fnslots * kMfBlock);  // every FP64 constant, then temps
}

void shadow_windows(int64_t max_steps, int64_t* win, int* nwin) {
  int64_t w = (max_steps + kMaxWin - 1) / kMaxWin;
  w = (w + kMfRange - 1) / kMfRange * kMfRange;
  if (w < kMfRange) w = kMfRange;
  int n = (int)((max_steps + w - 1) / w);
  *win = w;
  *nwin = n < 1 ? 1 : n;
}

void order_ranges(std::vector<Stage>& stages, std::vector<DevRange>& ranges, int64_t win, int nwin) {
  for (Stage& g : stages) {
    auto first = ranges.begin() + g.range_begin, last = ranges.begin() + g.range_end;
    std::stable_sort(first, last, [&](const DevRange& x, const DevRange& y) { return x.s_begin / win < y.s_begin / win; });
    g.win_range.assign(nwin + 1, g.range_end);
    g.win_range[0] = g.range_begin;
    int c = 0;
    for (int r = g.range_begin; r < g.range_end; ++r) {
      const int rc = (int)std::min<int64_t>(ranges[r].s_begin / win, nwin - 1);
      while (c < rc) g.win_range[++c] = r;
    }
    while (c < nwin) g.win_range[++c] = g.range_end;
  }
}

int launch_stages(drx_ctx* ctx, const KArgs& a, const std::vector<Stage>& stages, const DevTask* seed_tasks,
                  int finalize, bool timed, int64_t win, int nwin, bool mf_first) {
  hipStream_t s1 = ctx->stream, s2 = ctx->stream2, s3 = ctx->stream3;
  hipEvent_t* ev = ctx->ev;
  if (nwin < 1 || nwin > kMaxWin) return fail(ctx, DRX_EINVAL, "bad shadow window count %d", nwin);
  HIPCHK(ctx, hipEventRecord(ev[kEvFork], s1));
  HIPCHK(ctx, hipStreamWaitEvent(s2, ev[kEvFork], 0));
  HIPCHK(ctx, hipStreamWaitEvent(s3, ev[kEvFork], 0));
  hipStream_t smf = mf_first ? s1 : s2;
  // Of a pass in several windows, the first runs alone on the walk stream
  // ahead of the seeds: the shadow chain starts with it, and beside the seeds
  // and the walk it took three times as long (pcf2: 0.68 against 0.72 ms per
  // step). A pass of one window stays beside the seeds.
  const bool head = nwin > 1 || mf_first;
  auto mf_window = [&](int c) -> int {
    hipStream_t st = (c == 0 && head) ? s1 : smf;
    for (const Stage& g : stages) {
      const int r0 = g.win_range[c], r1 = g.win_range[c + 1];
      KArgs ma = a;
      ma.ranges += r0;
      HIPCHK(ctx, launch_mf(g.m, ma, r1 - r0, g.mf_lds, st));
    }
    HIPCHK(ctx, hipEventRecord(ev[kEvWin + c], st));
    if (c == 0 && st != smf) HIPCHK(ctx, hipStreamWaitEvent(smf, ev[kEvWin], 0));  // later windows after it
    return DRX_OK;
  };
  int rc = DRX_OK;
  // Enqueue order follows the critical paths: the first M_f window (which
  // the shadow chain waits for), then the seeds and the walk, the walk queued
  // before the later windows.
  if ((rc = mf_window(0))) return rc;
  for (const Stage& g : stages) {
    if (g.kind == WALK_INTERP || g.seed_end <= g.seed_begin) continue;
    KArgs fa = a;
    fa.tasks = seed_tasks + g.seed_begin;
    HIPCHK(ctx, launch_fd_seed(g.m, fa, g.seed_end - g.seed_begin, g.seed_kp, fd_seed_threads(g.m, g.seed_kp),
                               g.seed_lds, s1));
  }
  auto shadow_window = [&](int c) -> int {
    HIPCHK(ctx, hipStreamWaitEvent(s3, ev[kEvWin + c], 0));
    for (const Stage& g : stages) {
      KArgs sa = a;
      sa.spans += g.span_begin;
      HIPCHK(ctx, launch_shadow(g.m, sa, g.span_end - g.span_begin, (finalize && c == nwin - 1) ? 1 : 0,
                                (int64_t)c * win, c == nwin - 1 ? INT64_MAX : (int64_t)(c + 1) * win, s3));
    }
    return DRX_OK;
  };
  for (int c = 1; c < (mf_first ? nwin : 1); ++c)
    if ((rc = mf_window(c))) return rc;
  bool persist = false;
  for (const Stage& g : stages) persist = persist || g.persist > 0;
  if (persist) {
    if (!a.queue) return fail(ctx, DRX_EINVAL, "internal: persistent stages without task counters");
    HIPCHK(ctx, launch_fill_u32(a.queue, (int)stages.size(), 0u, s1));
  }
  if (timed) HIPCHK(ctx, hipEventRecord(ev[0], s1));
  for (size_t gi = 0; gi < stages.size(); ++gi) {
    const Stage& g = stages[gi];
    KArgs wa = a;
    int units = g.task_end - g.task_begin;
    if (g.wpg > 1) {  // task indices of the group table are plan-wide
      wa.gtask = a.gtask + (size_t)g.group_begin * g.wpg;
      wa.wpg = g.wpg;
      units = g.group_end - g.group_begin;
    } else {
      wa.tasks += g.task_begin;
      wa.gtask = nullptr;
      wa.wpg = 1;
    }
    HIPCHK(ctx, launch_walk(g.m, g.kind, wa, units, g.threads, g.lds, s1, g.persist,
                            g.persist > 0 ? a.queue + gi : nullptr));
  }
  if (timed) HIPCHK(ctx, hipEventRecord(ev[1], s1));
  for (int c = 0; c < nwin; ++c) {
    if (!mf_first && c + 1 < nwin && (rc = mf_window(c + 1))) return rc;
    if ((rc = shadow_window(c))) return rc;
  }
  // the M_f stream ends before stream3's last wait, so joining stream3 joins both
  HIPCHK(ctx, hipEventRecord(ev[kEvJoin], s3));
  for (const Stage& g : stages) {
    KArgs ca = a;
    ca.spans += g.span_begin;
    HIPCHK(ctx, launch_combine(g.m, ca, g.span_end - g.span_begin, s1));
  }
  if (timed) HIPCHK(ctx, hipEventRecord(ev[2], s1));
  return DRX_OK;
}

int join_stages(drx_ctx* ctx, bool timed) {
  HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev[kEvJoin], 0));
  if (timed) HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
  return DRX_OK;
}

// Inverse of a mod m for coprime a, m (extended Euclid).
static uint32_t inv_mod(uint32_t a, uint32_t m) {
  int64_t r0 = m, r1 = a % m, t0 = 0, t1 = 1;
  while (r1) {
    const int64_t q = r0 / r1, r2 = r0 - q * r1, t2 = t0 - q * t1;
    r0 = r1;
    r1 = r2;
    t0 = t1;
    t1 = t2;
  }
  return (uint32_t)(t0 < 0 ? t0 + m : t0);
}

// Garner's table inv[j][i] = p_j^{-1} mod p_i for the wavefront kernel's
// rows (at most 256 primes), built once for the first 256 primes of the
// context so the device copy never moves while plans hold it.
int ensure_inv_table(drx_ctx* ctx, int k) {
  if (k <= ctx->kinv) return DRX_OK;
  const int kall = std::min((int)ctx->primes.size(), kCrtMatrixMinK - 1);
  if (k > kall) return fail(ctx, DRX_EINVAL, "K=%d exceeds the %d primes of the Garner table", k, kall);
  std::vector<uint32_t> t((size_t)kall * kall, 0);
  for (int j = 0; j < kall; ++j)
    for (int i = 0; i < kall; ++i)
      if (i != j) t[(size_t)j * kall + i] = inv_mod(ctx->primes[j] % ctx->primes[i], ctx->primes[i]);
  if (ctx->d_inv.p) ctx->retired.push_back(ctx->d_inv.p);
  ctx->d_inv.p = nullptr;
  ctx->d_inv.cap = 0;
  HIPCHK(ctx, ctx->d_inv.ensure(t.size() * 4));
  HIPCHK(ctx, hipMemcpy(ctx->d_inv.p, t.data(), t.size() * 4, hipMemcpyHostToDevice));
  ctx->inv_table.swap(t);
  ctx->kinv = kall;
  return DRX_OK;
}

// w_i = (p_0 ... p_{i-1})^{-1} mod p_i for the wide Garner kernel, extended
// to the first k primes (O(i) per new entry); an outgrown device copy is
// retired, not freed, since plans hold its address.
int ensure_garner_w(drx_ctx* ctx, int k) {
  if (k > (int)ctx->primes.size()) return fail(ctx, DRX_EINVAL, "K=%d exceeds the %zu primes set", k, ctx->primes.size());
  const int have = (int)ctx->garner_w.size();
  if (k <= have && ctx->d_garner_w.p) return DRX_OK;
  // grow in steps so that a run of slightly wider rows does not rebuild it each time
  const int want = std::min<int>((int)ctx->primes.size(), std::max(k, std::max(2 * have, 256)));
  for (int i = have; i < want; ++i) {
    const uint64_t p = ctx->primes[i];
    uint64_t prod = 1 % p;
    for (int j = 0; j < i; ++j) prod = prod * (ctx->primes[j] % p) % p;
    ctx->garner_w.push_back((double)inv_mod((uint32_t)prod, (uint32_t)p));
  }
  HIPCHK(ctx, ctx->d_garner_w.ensure_keep(ctx->garner_w.size() * 8, ctx->retired));
  HIPCHK(ctx, hipMemcpy(ctx->d_garner_w.p, ctx->garner_w.data(), ctx->garner_w.size() * 8, hipMemcpyHostToDevice));
  return DRX_OK;
}

// Matrix-core CRT table of the first K primes: M = prod p_i as 32-bit limbs,
// w_i = (M / p_i)^-1 mod p_i, the balanced base-256 digits of every M / p_i
// (column i of BT, rows t; the top digit is the final carry), M and
// (M - 1) / 2 padded to L + 2 limbs. Built once per K; a prime change other
// than an extension drops the tables (their buffers are retired).
int ensure_crt_table(drx_ctx* ctx, int K, const CrtTable** out) {
  for (const auto& t : ctx->crt_tables)
    if (t->K == K) {
      *out = t.get();
      return DRX_OK;
    }
  if (K < 1 || K > (int)ctx->primes.size()) return fail(ctx, DRX_EINVAL, "K=%d exceeds the primes set", K);
  std::vector<uint32_t> M{1u};
  for (int i = 0; i < K; ++i) {
    uint64_t carry = 0;
    for (auto& limb : M) {
      const uint64_t x = (uint64_t)limb * ctx->primes[i] + carry;
      limb = (uint32_t)x;
      carry = x >> 32;
    }
    if (carry) M.push_back((uint32_t)carry);
  }
  std::unique_ptr<CrtTable> t(new CrtTable());
  t->K = K;
  t->L = (int)M.size();
  t->Kp = (K + 31) / 32 * 32;
  t->T = 4 * t->L + 1;
  t->Tp = (t->T + 3 + 127) / 128 * 128;  // the plane shifts reach column T + 2
  const int L = t->L, Kp = t->Kp;
  // kBtLead zero rows before row 0 (the GEMM reads rows t - b down to -3)
  std::vector<int8_t> bt((size_t)(kBtLead + t->Tp) * Kp, 0);
  int8_t* bt0 = bt.data() + (size_t)kBtLead * Kp;
  std::vector<double> w(K);
  std::vector<uint32_t> q(L);
  for (int i = 0; i < K; ++i) {
    const uint64_t p = ctx->primes[i];
    uint64_t rem = 0;
    for (int l = L - 1; l >= 0; --l) {  // q = M / p (exact)
      const uint64_t cur = (rem << 32) | M[l];
      q[l] = (uint32_t)(cur / p);
      rem = cur % p;
    }
    if (rem != 0) return fail(ctx, DRX_EINVAL, "internal: M not divisible by prime %d", i);
    uint64_t r = 0;
    for (int l = L - 1; l >= 0; --l) r = ((r << 32) | q[l]) % p;  // (M / p) mod p
    w[i] = (double)inv_mod((uint32_t)r, (uint32_t)p);
    int carry = 0;
    for (int d = 0; d < 4 * L; ++d) {
      int v = (int)((q[d / 4] >> (8 * (d % 4))) & 255u) + carry;
      carry = v >= 128 ? 1 : 0;
      bt0[(size_t)d * Kp + i] = (int8_t)(v - 256 * carry);
    }
    bt0[(size_t)(4 * L) * Kp + i] = (int8_t)carry;
  }
  std::vector<uint32_t> m(L + 2, 0u), h(L + 2, 0u);
  for (int l = 0; l < L; ++l) m[l] = M[l];
  for (int l = 0; l < L; ++l) h[l] = (M[l] >> 1) | (l + 1 < L ? (M[l + 1] << 31) : 0u);  // (M - 1) / 2, M odd
  HIPCHK(ctx, t->d_bt.ensure(bt.size()));
  HIPCHK(ctx, hipMemcpy(t->d_bt.p, bt.data(), bt.size(), hipMemcpyHostToDevice));
  HIPCHK(ctx, t->d_w.ensure(w.size() * 8));
  HIPCHK(ctx, hipMemcpy(t->d_w.p, w.data(), w.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(ctx, t->d_m.ensure(m.size() * 4));
  HIPCHK(ctx, hipMemcpy(t->d_m.p, m.data(), m.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(ctx, t->d_h.ensure(h.size() * 4));
  HIPCHK(ctx, hipMemcpy(t->d_h.p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
  *out = t.get();
  ctx->crt_tables.push_back(std::move(t));
  return DRX_OK;
}

void retire_crt_tables(drx_ctx* ctx) {
  for (auto& t : ctx->crt_tables)
    for (DBuf* d : {&t->d_bt, &t->d_w, &t->d_m, &t->d_h})
      if (d->p) {
        ctx->retired.push_back(d->p);
        d->p = nullptr;
        d->cap = 0;
      }
  ctx->crt_tables.clear();
}

int axis_tables(drx_ctx* ctx, const int64_t* q, int dim, int K, int64_t depth, double* axr, int64_t* axi,
                double* axd) {
  for (int ax = 0; ax < dim; ++ax) {
    const __int128 sn = q[4 * ax], sd = q[4 * ax + 1], tn = q[4 * ax + 2], td = q[4 * ax + 3];
    const __int128 c0 = sn * td, c1 = tn * sd, den = sd * td;
    if (den == 0) return fail(ctx, DRX_EINVAL, "axis %d: zero denominator", ax);
    const __int128 lim = (__int128)1 << 62;
    const __int128 last = c0 + (__int128)depth * c1;
    if (c0 > lim || c0 < -lim || c1 > lim || c1 < -lim || den > lim || den < -lim || last > lim || last < -lim)
      return fail(ctx, DRX_ERANGE, "axis %d: shift/trajectory too large for int64", ax);
    double* base = axr + (size_t)ax * 3 * K;
    for (int k = 0; k < K; ++k) {
      const uint32_t p = ctx->primes[k];
      const uint32_t dr = mod_i128(den, p);
      base[k] = balanced(mod_i128(c0, p), p);
      base[K + k] = balanced(mod_i128(c1, p), p);
      // inverse of the combined denominator; 0 when p divides it (Fermat)
      base[2 * K + k] = balanced(dr == 1 ? 1u : powmod(dr, p - 2, p), p);
    }
    axi[2 * ax] = (int64_t)c0;
    axi[2 * ax + 1] = (int64_t)c1;
    axd[ax] = (double)(int64_t)den;
  }
  return DRX_OK;
}

int init_tables(drx_ctx* ctx, const drx_walk& w, int m, const uint32_t* init_res_pool, const double* init_f_pool,
                double* res, double* f) {
  const int K = w.K;
  for (int e = 0; e < m * m; ++e) {
    f[e] = 0.0;
    for (int k = 0; k < K; ++k) res[(size_t)e * K + k] = 0.0;
  }
  f[m * m] = 1.0;      // log_scale = ln(1.0) + 0 * ln 2
  f[m * m + 1] = 0.0;
  if (w.init_kind == DRX_INIT_GIVEN) {
    for (int e = 0; e < m * m; ++e) {
      for (int k = 0; k < K; ++k) {
        const uint32_t v = init_res_pool[w.init_off + (int64_t)e * K + k];
        if (v >= ctx->primes[k]) return fail(ctx, DRX_EINVAL, "initial residue not canonical");
        res[(size_t)e * K + k] = (double)v;
      }
      f[e] = init_f_pool[w.init_f_off + e];
    }
    return DRX_OK;
  }
  for (int d = 0; d < m; ++d) {
    for (int k = 0; k < K; ++k) res[(size_t)(d * m + d) * K + k] = 1.0;
    f[d * m + d] = 1.0;
  }
  if (w.init_kind == DRX_INIT_PCF) {
    for (int k = 0; k < K; ++k) res[(size_t)1 * K + k] = (double)mod_i64(w.a0, ctx->primes[k]);
    f[1] = (double)w.a0;
  }
  return DRX_OK;
}
}  // namespace drx

// ======================================================================= plan
struct drx_plan {
  drx_ctx* ctx = nullptr;
  drx_batch b{};
  drx_launch_cfg cfg{};
  uint64_t epoch = 0;  // context epoch the plan was built in
  int n = 0, kmax = 0, ld = 0;
  std::vector<Stage> stages;
  int64_t win = 1;  // FP64 chain window (steps) and count
  int nwin = 1;
  std::vector<int> walk_order;  // span index -> walk
  DBuf d_walks, d_tasks, d_seeds, d_spans, d_ranges, d_res0, d_fst0, d_res, d_fst, d_axr, d_axi, d_axd, d_cp, d_mf,
      d_fdt;
  DBuf d_fout, d_rout, d_pout, d_pfull, d_limbs, d_neg, d_topk, d_topk_ws, d_rowk;
  DBuf d_queue;  // persistent mode: one task counter per stage
  DBuf d_gtask;  // walks per workgroup: the stages' task groups
  bool mf_first = false;  // the M_f pass before the walk (see launch_stages)
  // every output sink is page-locked host memory: the outputs that need only
  // the combined residues are copied while the shadow chain still runs
  bool early_out = false;
  KArgs args{};
  GarnerArgs gargs{};   // rows of at most 256 primes: the wavefront Garner kernel
  GarnerArgs wargs{};   // wider rows outside the matrix path: the wide Garner kernel
  // matrix-core CRT: one group per prime count; the groups run one after the
  // other on one stream and share one scratch (digit planes, plane products, k)
  struct CrtGroup {
    DBuf d_rows;
    CrtArgs args{};
  };
  std::vector<std::unique_ptr<CrtGroup>> crt_groups;
  DBuf d_grows, d_wrows, d_crt_a, d_crt_c, d_crt_k;
  int ntask = 0;
  double residue_steps = 0, mm_total = 0, ma_total = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<drx_plan*> plain;  // PLAIN mode: one sub-plan per walk
  // state-vector walks: residue groups, FP64 groups, walk order
  bool vec = false;
  DBuf d_vgroups, d_fgroups, d_vorder;
  int nvg = 0, nfg = 0, vthreads = 0, fthreads = 0;
  size_t vlds = 0, flds = 0;
  ~drx_plan() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (DBuf* d : {&d_walks, &d_tasks, &d_seeds, &d_spans, &d_ranges, &d_res0, &d_fst0, &d_res, &d_fst, &d_axr, &d_axi,
                    &d_axd, &d_cp, &d_mf, &d_fdt, &d_fout, &d_rout, &d_pout, &d_pfull, &d_limbs, &d_neg, &d_topk,
                    &d_topk_ws, &d_rowk, &d_vgroups, &d_fgroups, &d_vorder, &d_grows, &d_wrows, &d_crt_a,
                    &d_crt_c, &d_crt_k, &d_queue, &d_gtask})
      d->release();
    for (auto& g : crt_groups) g->d_rows.release();
    for (drx_plan* p : plain) delete p;
  }
};

namespace {

template <typename T>
int upload(drx_ctx* ctx, DBuf& d, const std::vector<T>& h) {
  HIPCHK(ctx, d.ensure(std::max<size_t>(1, h.size()) * sizeof(T)));
  if (!h.empty()) HIPCHK(ctx, hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return DRX_OK;
}

// CRT of the plan's p and q rows: the wavefront Garner kernel for rows of at
// most 256 primes, the matrix-core CRT for prime counts with enough rows
// (kCrtMatrixMinRows, up to kCrtMatrixMaxK primes), the wide Garner kernel
// for the other wide rows.
int launch_crt(drx_plan* pl, hipStream_t s) {
  drx_ctx* ctx = pl->ctx;
  if (pl->gargs.n > 0) HIPCHK(ctx, launch_garner(pl->gargs, s));
  if (pl->wargs.n > 0) HIPCHK(ctx, launch_garner_wide(pl->wargs, s));
  for (const auto& g : pl->crt_groups) HIPCHK(ctx, launch_crt_matrix(g->args, s));
  return DRX_OK;
}

// The CRT path of rows of K primes when `nrows` rows share that K.
enum CrtPath { CRT_WAVE = 0, CRT_WIDE = 1, CRT_MATRIX = 2 };
static int crt_path(int K, size_t nrows) {
  if (K < kCrtMatrixMinK) return CRT_WAVE;
  return (K <= kCrtMatrixMaxK && nrows >= (size_t)kCrtMatrixMinRows) ? CRT_MATRIX : CRT_WIDE;
}

// Matrix-core CRT arguments of one group (rows `rows`, scratch shared).
static void crt_matrix_args(drx_ctx* ctx, const CrtTable* t, int n, int npad, int32_t* rows, const int64_t* res,
                            uint64_t stride, int8_t* A, int32_t* C, int32_t* kq, double* fd, uint32_t* limbs,
                            uint8_t* neg, CrtArgs& c) {
  c.res = res;
  c.res_stride = stride;
  c.rows = rows;
  c.n = n;
  c.npad = npad;
  c.K = t->K;
  c.Kp = t->Kp;
  c.T = t->T;
  c.Tp = t->Tp;
  c.L = t->L;
  c.prime_p = ctx->d_prime_p.as<double>();
  c.prime_pinv = ctx->d_prime_pinv.as<double>();
  c.w = t->d_w.as<double>();
  c.bt = t->d_bt.as<int8_t>() + (size_t)kBtLead * t->Kp;
  c.m = t->d_m.as<uint32_t>();
  c.h = t->d_h.as<uint32_t>();
  c.A = A;
  c.C = C;
  c.kq = kq;
  c.fd = fd;
  c.limbs = limbs;
  c.neg = neg;
}

// Split the CRT rows (2 per walk) by path; runs after the outputs are
// allocated and GarnerArgs' common fields are set.
int setup_crt(drx_ctx* ctx, drx_plan* pl) {
  const drx_batch& b = pl->b;
  pl->crt_groups.clear();
  GarnerArgs& g = pl->gargs;
  GarnerArgs& wg = pl->wargs;
  wg = g;
  wg.n = 0;
  if (!b.crt_limbs) {
    g.n = 0;
    return DRX_OK;
  }
  std::vector<std::pair<int, std::vector<int32_t>>> byk;  // rows by prime count
  for (int i = 0; i < pl->n; ++i) {
    const int K = b.walks[i].K;
    auto it = std::find_if(byk.begin(), byk.end(), [&](const std::pair<int, std::vector<int32_t>>& e) {
      return e.first == K;
    });
    if (it == byk.end()) {
      byk.push_back({K, {}});
      it = byk.end() - 1;
    }
    it->second.push_back(2 * i);
    it->second.push_back(2 * i + 1);
  }
  std::vector<int32_t> grows, wrows;
  std::vector<std::pair<int, std::vector<int32_t>>> wide;
  int gk = 1, wk = 1;
  for (auto& e : byk) {
    switch (crt_path(e.first, e.second.size())) {
      case CRT_WAVE:
        grows.insert(grows.end(), e.second.begin(), e.second.end());
        gk = std::max(gk, e.first);
        break;
      case CRT_WIDE:
        wrows.insert(wrows.end(), e.second.begin(), e.second.end());
        wk = std::max(wk, e.first);
        break;
      default:
        wide.push_back(std::move(e));
    }
  }
  int rc;
  const bool all_wave = grows.size() == (size_t)2 * pl->n;
  g.n = (int)grows.size();
  g.K = gk;
  g.rows = nullptr;
  if (!grows.empty() && !all_wave) {
    if ((rc = upload(ctx, pl->d_grows, grows))) return rc;
    g.rows = pl->d_grows.as<int32_t>();
  }
  if (!grows.empty()) {
    if ((rc = ensure_inv_table(ctx, gk))) return rc;
    g.inv_table = ctx->d_inv.as<uint32_t>();
    g.inv_ld = (uint64_t)ctx->kinv;
  }
  if (!wrows.empty()) {
    if ((rc = ensure_garner_w(ctx, wk))) return rc;
    if ((rc = upload(ctx, pl->d_wrows, wrows))) return rc;
    wg.rows = pl->d_wrows.as<int32_t>();
    wg.n = (int)wrows.size();
    wg.K = wk;
    wg.w = ctx->d_garner_w.as<double>();
  }
  size_t a_bytes = 0, c_bytes = 0, k_bytes = 0;
  std::vector<const CrtTable*> tabs;
  for (auto& e : wide) {
    const CrtTable* t = nullptr;
    if ((rc = ensure_crt_table(ctx, e.first, &t))) return rc;
    if ((size_t)t->L + 2 > (size_t)pl->ld) return fail(ctx, DRX_EINVAL, "internal: CRT rows narrower than M");
    const size_t npad = (e.second.size() + 127) / 128 * 128;
    a_bytes = std::max(a_bytes, (size_t)4 * npad * t->Kp);
    c_bytes = std::max(c_bytes, (size_t)t->Tp * npad * 4);
    k_bytes = std::max(k_bytes, npad * 12);
    tabs.push_back(t);
  }
  if (!wide.empty()) {
    HIPCHK(ctx, pl->d_crt_a.ensure(a_bytes));
    HIPCHK(ctx, pl->d_crt_c.ensure(c_bytes));
    HIPCHK(ctx, pl->d_crt_k.ensure(k_bytes));
  }
  for (size_t gi = 0; gi < wide.size(); ++gi) {
    const auto& e = wide[gi];
    std::unique_ptr<drx_plan::CrtGroup> grp(new drx_plan::CrtGroup());
    const int ng = (int)e.second.size(), npad = (ng + 127) / 128 * 128;
    if ((rc = upload(ctx, grp->d_rows, e.second))) return rc;
    crt_matrix_args(ctx, tabs[gi], ng, npad, grp->d_rows.as<int32_t>(), pl->d_rout.as<int64_t>(), (uint64_t)pl->ld,
                    pl->d_crt_a.as<int8_t>(), pl->d_crt_c.as<int32_t>(), pl->d_crt_k.as<int32_t>(),
                    reinterpret_cast<double*>(pl->d_crt_k.as<char>() + (size_t)npad * 4), pl->d_limbs.as<uint32_t>(),
                    pl->d_neg.as<uint8_t>(), grp->args);
    pl->crt_groups.push_back(std::move(grp));
  }
  return DRX_OK;
}

// Vector walks: the residue kernel on the library's stream and the FP64
// shadow on stream2 side by side; CRT after the residues, top-k behind the
// shadow on stream2, then the join.
int launch_vec_sequence(drx_plan* pl, bool timed) {
  drx_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream, s2 = ctx->stream2;
  hipEvent_t* ev = ctx->ev;
  int rc;
  HIPCHK(ctx, hipEventRecord(ev[kEvFork], s));
  HIPCHK(ctx, hipStreamWaitEvent(s2, ev[kEvFork], 0));
  HIPCHK(ctx, launch_vec_fp(pl->args, pl->nfg, pl->fthreads, pl->flds, s2));
  if (pl->b.top_k > 0)
    HIPCHK(ctx, launch_topk(pl->d_fout.as<double>() + 3, 6, pl->n, pl->b.top_k, pl->d_topk.as<int32_t>(),
                            pl->d_topk_ws.p, s2));
  HIPCHK(ctx, hipEventRecord(ev[kEvJoin], s2));
  if (timed) HIPCHK(ctx, hipEventRecord(ev[0], s));
  HIPCHK(ctx, launch_vec_walk(pl->args, pl->nvg, pl->vthreads, pl->vlds, s));
  if (timed) {
    HIPCHK(ctx, hipEventRecord(ev[1], s));
    HIPCHK(ctx, hipEventRecord(ev[2], s));
  }
  if (pl->b.crt_limbs && (rc = launch_crt(pl, s))) return rc;
  HIPCHK(ctx, hipStreamWaitEvent(s, ev[kEvJoin], 0));
  if (timed) {
    HIPCHK(ctx, hipEventRecord(ev[3], s));
    HIPCHK(ctx, hipEventRecord(ev[4], s));
  }
  return DRX_OK;
}

int plan_fetch(drx_plan* pl);

// The launch sequence of a plan; with `fetch`, also the copy of every output
// to the caller's sinks, ending on ctx->stream.
int launch_sequence(drx_plan* pl, bool timed, bool fetch = false) {
  if (pl->vec) {
    const int rc = launch_vec_sequence(pl, timed);
    return rc || !fetch ? rc : plan_fetch(pl);
  }
  drx_ctx* ctx = pl->ctx;
  hipStream_t s = ctx->stream;
  int rc = launch_stages(ctx, pl->args, pl->stages, pl->d_seeds.as<DevTask>(), 1, timed, pl->win, pl->nwin,
                         pl->mf_first);
  if (rc) return rc;
  if (pl->b.top_k > 0) {  // on the shadow's stream, right behind the delta_float it ranks
    HIPCHK(ctx, launch_topk(pl->d_fout.as<double>() + 3, 6, pl->n, pl->b.top_k, pl->d_topk.as<int32_t>(),
                            pl->d_topk_ws.p, ctx->stream3));
    HIPCHK(ctx, hipEventRecord(ctx->ev[kEvJoin], ctx->stream3));
  }
  // residue outputs and CRT need only the combined residues: they run while
  // the shadow chain is still going; the final P_f copy needs the shadow
  const bool ffull = pl->args.ffull != nullptr;
  const drx_batch& b = pl->b;
  // With page-locked sinks the residues and the CRT limbs leave on stream2
  // (its M_f windows are queued before) as soon as they are final, beside
  // the rest of the shadow chain.
  const bool early = fetch && pl->early_out && !ffull;
  hipStream_t s2 = ctx->stream2;
  hipEvent_t* ev = ctx->ev;
  const size_t n = (size_t)pl->n, ld = (size_t)pl->ld;
  if (!ffull) HIPCHK(ctx, launch_finalize(pl->args, (int)pl->walk_order.size(), b.full != nullptr, s));
  if (early) {
    HIPCHK(ctx, hipEventRecord(ev[kEvOut], s));
    HIPCHK(ctx, hipStreamWaitEvent(s2, ev[kEvOut], 0));
    if (b.residues) HIPCHK(ctx, hipMemcpyAsync(b.residues, pl->d_rout.p, n * 2 * ld * 8, hipMemcpyDeviceToHost, s2));
    if (b.full) HIPCHK(ctx, hipMemcpyAsync(b.full, pl->d_pout.p, n * 16 * ld * 8, hipMemcpyDeviceToHost, s2));
  }
  if (!ffull && b.crt_limbs && (rc = launch_crt(pl, s))) return rc;
  if (early && b.crt_limbs) {
    HIPCHK(ctx, hipEventRecord(ev[kEvLimbs], s));
    HIPCHK(ctx, hipStreamWaitEvent(s2, ev[kEvLimbs], 0));
    HIPCHK(ctx, hipMemcpyAsync(b.crt_limbs, pl->d_limbs.p, n * 2 * ld * 4, hipMemcpyDeviceToHost, s2));
    if (b.crt_neg) HIPCHK(ctx, hipMemcpyAsync(b.crt_neg, pl->d_neg.p, n * 2, hipMemcpyDeviceToHost, s2));
  }
  if (early) HIPCHK(ctx, hipEventRecord(ev[kEvCopied], s2));
  if ((rc = join_stages(ctx, timed))) return rc;
  if (ffull) HIPCHK(ctx, launch_finalize(pl->args, (int)pl->walk_order.size(), b.full != nullptr, s));
  if (ffull && b.crt_limbs && (rc = launch_crt(pl, s))) return rc;
  if (timed) HIPCHK(ctx, hipEventRecord(ev[4], s));
  if (!fetch) return DRX_OK;
  if (!early) return plan_fetch(pl);
  HIPCHK(ctx, hipStreamWaitEvent(s, ev[kEvCopied], 0));
  if (b.out) HIPCHK(ctx, hipMemcpyAsync(b.out, pl->d_fout.p, n * 6 * 8, hipMemcpyDeviceToHost, s));
  if (b.top_k > 0 && b.topk_idx)
    HIPCHK(ctx, hipMemcpyAsync(b.topk_idx, pl->d_topk.p, (size_t)b.top_k * 4, hipMemcpyDeviceToHost, s));
  return DRX_OK;
}

// The launch sequence (without the copies back) captured once as a graph
// that every run replays.
int capture_plan(drx_plan* pl) {
  drx_ctx* ctx = pl->ctx;
  HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = launch_sequence(pl, false);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(ctx, DRX_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  pl->graph = g;
  HIPCHK(ctx, hipGraphInstantiate(&pl->gexec, g, nullptr, nullptr, 0));
  return DRX_OK;
}

// Page-locked host memory (or no sink at all).
bool page_locked(const void* p) {
  if (!p) return true;
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();  // pageable memory: leave no error behind
    return false;
  }
  return at.type == hipMemoryTypeHost;
}

bool sinks_page_locked(const drx_batch& b) {
  return page_locked(b.out) && page_locked(b.residues) && page_locked(b.full) && page_locked(b.full_f) &&
         page_locked(b.crt_limbs) && page_locked(b.crt_neg) && page_locked(b.topk_idx);
}

// Kernel arguments shared by both plan kinds: device pools and outputs.
void set_common_args(drx_ctx* ctx, drx_plan* pl) {
  const drx_batch& b = pl->b;
  pl->early_out = sinks_page_locked(b);
  KArgs& a = pl->args;
  a.walks = pl->d_walks.as<DevWalk>();
  a.progs = ctx->d_progs.as<DevProgram>();
  a.invs = ctx->d_invs.as<double>();
  a.code = ctx->d_code.as<uint32_t>();
  a.cres = ctx->d_cres.as<double>();
  a.cflt = ctx->d_cflt.as<double>();
  a.prime_p = ctx->d_prime_p.as<double>();
  a.prime_pinv = ctx->d_prime_pinv.as<double>();
  a.prime_e = ctx->d_prime_e.as<uint32_t>();
  a.axr = pl->d_axr.as<double>();
  a.axi = pl->d_axi.as<int64_t>();
  a.axd = pl->d_axd.as<double>();
  a.res_in = pl->d_res0.as<double>();
  a.fst_in = pl->d_fst0.as<double>();
  a.fout = pl->d_fout.as<double>();
  a.rout = pl->d_rout.as<int64_t>();
  a.pout = b.full ? pl->d_pout.as<int64_t>() : nullptr;
  a.fcompact = nullptr;
  a.ffull = b.full_f ? pl->d_pfull.as<double>() : nullptr;
  a.out_ld = (uint64_t)pl->ld;
  GarnerArgs& g = pl->gargs;
  g.res = pl->d_rout.as<int64_t>();
  g.res_stride = (uint64_t)pl->ld;
  g.n = b.crt_limbs ? 2 * pl->n : 0;
  g.K = pl->kmax;
  g.row_k = b.crt_limbs ? pl->d_rowk.as<int32_t>() : nullptr;
  g.prime_p = ctx->d_prime_p.as<double>();
  g.prime_pinv = ctx->d_prime_pinv.as<double>();
  g.prime_u = ctx->d_prime_u.as<uint32_t>();
  g.inv_table = ctx->d_inv.as<uint32_t>();
  g.inv_ld = (uint64_t)ctx->kinv;
  g.limbs = pl->d_limbs.as<uint32_t>();
  g.neg = pl->d_neg.as<uint8_t>();
}

// Output buffers shared by both plan kinds.
int alloc_outputs(drx_ctx* ctx, drx_plan* pl) {
  const drx_batch& b = pl->b;
  const int n = pl->n;
  HIPCHK(ctx, pl->d_fout.ensure((size_t)std::max(1, n) * 6 * 8));
  HIPCHK(ctx, pl->d_rout.ensure((size_t)std::max(1, n) * 2 * pl->ld * 8));
  if (b.full) HIPCHK(ctx, pl->d_pout.ensure((size_t)std::max(1, n) * 16 * pl->ld * 8));
  if (b.full_f) HIPCHK(ctx, pl->d_pfull.ensure((size_t)std::max(1, n) * 16 * 8));
  if (b.top_k > 0) {
    if (b.top_k > n) return fail(ctx, DRX_EINVAL, "top_k=%d > n=%d", b.top_k, n);
    HIPCHK(ctx, pl->d_topk.ensure((size_t)b.top_k * 4));
    HIPCHK(ctx, pl->d_topk_ws.ensure(std::max<size_t>(16, topk_workspace_bytes(n, b.top_k))));
  }
  if (b.crt_limbs) {
    int rc;
    std::vector<int32_t> rowk(2 * (size_t)n);
    for (int i = 0; i < n; ++i) rowk[2 * i] = rowk[2 * i + 1] = b.walks[i].K;
    if ((rc = upload(ctx, pl->d_rowk, rowk))) return rc;
    HIPCHK(ctx, pl->d_limbs.ensure((size_t)std::max(1, n) * 2 * pl->ld * 4));
    HIPCHK(ctx, pl->d_neg.ensure((size_t)std::max(1, n) * 2));
  }
  return DRX_OK;
}

// State-vector walks (cmf_walk.py:445-531). Walks are ordered by (program,
// depth, K); consecutive walks sharing them form a residue group of
// T / KP walks (one lane per walk and prime) and an FP64 group of up to
// 256 walks (one lane per walk), so each workgroup interprets one program for
// one trip count.
int plan_build_vec(drx_ctx* ctx, drx_plan* pl) {
  const drx_batch& b = pl->b;
  const int n = pl->n;
  pl->vec = true;
  const int thr = pl->cfg.threads > 0 ? pl->cfg.threads : 256;
  if ((thr & (thr - 1)) || thr < 32 || thr > 256)
    return fail(ctx, DRX_EINVAL, "threads=%d must be a power of two in [32, 256]", thr);
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    const drx_walk &wx = b.walks[x], &wy = b.walks[y];
    if (wx.prog != wy.prog) return wx.prog < wy.prog;
    if (wx.depth != wy.depth) return wx.depth > wy.depth;
    return wx.K < wy.K;
  });
  std::vector<DevWalk> hw(n);
  std::vector<double> res0, fst0, axr, axd;
  std::vector<int64_t> axi;
  double rsteps = 0, mmt = 0, mat = 0;
  int kpmax = 1;
  for (int i = 0; i < n; ++i) {
    const drx_walk& w = b.walks[i];
    const HostProgram& hp = ctx->progs[w.prog];
    const int m = (int)hp.dev.m, dim = (int)hp.dev.dim, K = w.K;
    if (w.init_kind != DRX_INIT_GIVEN) return fail(ctx, DRX_EINVAL, "walk %d: a vector walk needs a given state", i);
    if (w.out_p < 0 || w.out_p >= m || w.out_q < 0 || w.out_q >= m)
      return fail(ctx, DRX_EINVAL, "walk %d: state entries %d, %d outside [0, %d)", i, w.out_p, w.out_q, m);
    DevWalk& dw = hw[i];
    dw.prog = w.prog;
    dw.K = K;
    dw.m = m;
    dw.dim = dim;
    dw.pe = w.out_p;
    dw.qe = w.out_q;
    dw.target = w.target;
    dw.res_off = res0.size();
    dw.fst_off = fst0.size();
    dw.axr_off = axr.size();
    dw.axi_off = axi.size();
    dw.axd_off = axd.size();
    for (int j = 0; j < m; ++j) {
      for (int k = 0; k < K; ++k) {
        const uint32_t v = b.init_res_pool[w.init_off + (int64_t)j * K + k];
        if (v >= ctx->primes[k]) return fail(ctx, DRX_EINVAL, "walk %d: initial residue not canonical", i);
        res0.push_back((double)v);
      }
      fst0.push_back(b.init_f_pool[w.init_f_off + j]);
    }
    const size_t a0 = axr.size(), i0 = axi.size(), d0 = axd.size();
    axr.resize(a0 + (size_t)dim * 3 * K);
    axi.resize(i0 + (size_t)dim * 2);
    axd.resize(d0 + (size_t)dim);
    if (dim > 0) {
      const int rc0 = axis_tables(ctx, b.axis_pool + w.axis_off, dim, K, w.depth, axr.data() + a0, axi.data() + i0,
                                  axd.data() + d0);
      if (rc0) return fail(ctx, rc0, "walk %d: %s", i, std::string(ctx->err).c_str());
    }
    kpmax = std::max<int>(kpmax, pow2_ceil(K));
    rsteps += (double)w.depth * K;
    mmt += (double)w.depth * K * hp.mm_ops;
    mat += (double)w.depth * K * hp.ma_ops;
  }
  if (kpmax > thr) return fail(ctx, DRX_EINVAL, "threads=%d below the %d lanes of one walk's primes", thr, kpmax);
  // groups
  std::vector<DevVGroup> vg, fg;
  std::vector<int32_t> vorder(order.begin(), order.end());
  for (int a0 = 0; a0 < n;) {
    const drx_walk& w = b.walks[order[a0]];
    int e = a0;
    while (e < n && b.walks[order[e]].prog == w.prog && b.walks[order[e]].depth == w.depth) ++e;
    for (int f = a0; f < e; f += 256) {  // FP64 groups: one lane per walk
      DevVGroup g{};
      g.first = f;
      g.n = std::min(256, e - f);
      g.prog = w.prog;
      g.K = 0;
      g.depth = w.depth;
      fg.push_back(g);
    }
    for (int f = a0; f < e;) {  // residue groups: equal K, thr / KP walks
      const int K = b.walks[order[f]].K;
      const int per = thr / (int)pow2_ceil(K);
      int c = f;
      while (c < e && c - f < per && b.walks[order[c]].K == K) ++c;
      DevVGroup g{};
      g.first = f;
      g.n = c - f;
      g.prog = w.prog;
      g.K = K;
      g.depth = w.depth;
      vg.push_back(g);
      f = c;
    }
    a0 = e;
  }
  // LDS: residue groups (constant rows by prime; spill, state A/B and
  // increment rows by thread); FP64 groups (constants, temps)
  const size_t lds_limit = device_lds_limit(ctx);
  size_t vlds = 8, flds = 8;
  for (const DevVGroup& g : vg) {
    const DevProgram& dp = ctx->progs[g.prog].dev;
    const size_t nl = lds_const_rows(dp);
    vlds = std::max(vlds, 8 * (nl * pow2_ceil(g.K) + ((size_t)dp.nslots + 2 * dp.m + dp.dim) * (size_t)thr));
  }
  for (const DevVGroup& g : fg) {
    const DevProgram& dp = ctx->progs[g.prog].dev;
    flds = std::max(flds, 8 * ((size_t)dp.nconst + (size_t)dp.fnslots * 256));
  }
  if (vlds > lds_limit || flds > lds_limit)
    return fail(ctx, DRX_ERANGE, "vector walks need %zu / %zu bytes of LDS (limit %zu)", vlds, flds, lds_limit);
  pl->nvg = (int)vg.size();
  pl->nfg = (int)fg.size();
  pl->vthreads = thr;
  pl->fthreads = 256;
  pl->vlds = vlds;
  pl->flds = flds;
  pl->residue_steps = rsteps;
  pl->mm_total = mmt;
  pl->ma_total = mat;
  pl->ntask = pl->nvg;
  int rc;
  if ((rc = upload(ctx, pl->d_walks, hw)) || (rc = upload(ctx, pl->d_res0, res0)) || (rc = upload(ctx, pl->d_fst0, fst0)) ||
      (rc = upload(ctx, pl->d_axr, axr)) || (rc = upload(ctx, pl->d_axi, axi)) || (rc = upload(ctx, pl->d_axd, axd)) ||
      (rc = upload(ctx, pl->d_vgroups, vg)) || (rc = upload(ctx, pl->d_fgroups, fg)) ||
      (rc = upload(ctx, pl->d_vorder, vorder)))
    return rc;
  if ((rc = alloc_outputs(ctx, pl))) return rc;
  if ((rc = upload_programs(ctx))) return rc;
  set_common_args(ctx, pl);
  if ((rc = setup_crt(ctx, pl))) return rc;
  pl->args.vgroups = pl->d_vgroups.as<DevVGroup>();
  pl->args.fgroups = pl->d_fgroups.as<DevVGroup>();
  pl->args.vorder = pl->d_vorder.as<int32_t>();
  for (int i = 0; i < n; ++i) pl->walk_order.push_back(i);
  if (pl->cfg.use_graph && (rc = capture_plan(pl))) return rc;
  return DRX_OK;
}

// The walk kernel's task array is laid out per walk in step order; stages
// hold one matrix order and walk-kernel family each; within a stage walks run
// longest first.
int plan_build(drx_ctx* ctx, drx_plan* pl) {
  const drx_batch& b = pl->b;
  const int n = b.n;
  const int np = (int)ctx->progs.size();
  const int K_ctx = (int)ctx->primes.size();
  int kmax = 1, nvec = 0;
  for (int i = 0; i < n; ++i) {
    const drx_walk& w = b.walks[i];
    if (w.prog < 0 || w.prog >= np) return fail(ctx, DRX_EINVAL, "walk %d: bad program id %d", i, w.prog);
    if (w.K < 1 || w.K > K_ctx)
      return fail(ctx, DRX_EINVAL, "walk %d: K=%d outside [1, %d] (primes set)", i, w.K, K_ctx);
    if (w.K > ctx->kwalk)
      return fail(ctx, DRX_ERANGE, "walk %d: K=%d but only the first %d primes lie in (2^30, 2^31), as walks need",
                  i, w.K, ctx->kwalk);
    const HostProgram& hp = ctx->progs[w.prog];
    if (w.K > (int)hp.dev.kcap)
      return fail(ctx, DRX_EINVAL, "walk %d: K=%d exceeds the program's constant table (%u)", i, w.K, hp.dev.kcap);
    if (w.depth < 0 || w.depth > (int64_t)1 << 30)
      return fail(ctx, DRX_EINVAL, "walk %d: depth %lld outside [0, 2^30]", i, (long long)w.depth);
    if (w.init_kind < DRX_INIT_PCF || w.init_kind > DRX_INIT_GIVEN)
      return fail(ctx, DRX_EINVAL, "walk %d: bad init kind", i);
    if (w.init_kind == DRX_INIT_PCF && hp.dev.m != 2)
      return fail(ctx, DRX_EINVAL, "walk %d: PCF initial matrix needs a 2x2 program", i);
    if (w.init_kind == DRX_INIT_GIVEN && (!b.init_res_pool || !b.init_f_pool))
      return fail(ctx, DRX_EINVAL, "walk %d: DRX_INIT_GIVEN without init pools", i);
    if (hp.dev.dim > 0 && !b.axis_pool) return fail(ctx, DRX_EINVAL, "walk %d: no axis pool", i);
    if (w.mode == DRX_WALK_VECTOR && pow2_ceil(w.K) > 256)
      return fail(ctx, DRX_ERANGE, "walk %d: K=%d above the 256 primes of a vector walk", i, w.K);
    if (w.mode != DRX_WALK_MATRIX && w.mode != DRX_WALK_VECTOR) return fail(ctx, DRX_EINVAL, "walk %d: bad mode", i);
    if (w.mode != hp.mode)
      return fail(ctx, DRX_EINVAL, "walk %d: program %d was uploaded for the other walk mode", i, w.prog);
    nvec += w.mode == DRX_WALK_VECTOR;
    kmax = std::max(kmax, w.K);
  }
  if (b.res_stride < kmax) return fail(ctx, DRX_EINVAL, "res_stride %d < max K %d", b.res_stride, kmax);
  pl->n = n;
  pl->kmax = kmax;
  pl->ld = b.res_stride;
  if (nvec > 0 && nvec < n) return fail(ctx, DRX_EINVAL, "a batch holds walks of one mode (matrix or vector)");
  if (nvec > 0) return plan_build_vec(ctx, pl);

  const int chunk_cfg = pl->cfg.chunk_steps > 0 ? pl->cfg.chunk_steps : 4096;
  const int thr_cfg = pl->cfg.threads > 0 ? pl->cfg.threads : 256;
  if ((thr_cfg & (thr_cfg - 1)) || thr_cfg < 32 || thr_cfg > 256)
    return fail(ctx, DRX_EINVAL, "threads=%d must be a power of two in [32, 256]", thr_cfg);
  const int slice = pl->cfg.slice > 0 ? pl->cfg.slice : kSliceK;
  if ((slice & (slice - 1)) || slice < 16 || slice > kSliceK)
    return fail(ctx, DRX_EINVAL, "slice=%d must be a power of two in [16, %d]", slice, kSliceK);
  const int wpg_cfg = std::max(1, pl->cfg.walks_per_wg);
  if (wpg_cfg > 256 || (wpg_cfg & (wpg_cfg - 1)))
    return fail(ctx, DRX_EINVAL, "walks_per_wg=%d must be a power of two up to 256", wpg_cfg);

  // order: by matrix order and walk-kernel family, then cost descending
  // (longest first), then the companion-PCF shape (equal shapes adjacent)
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  auto cost = [&](int i) {
    const drx_walk& w = b.walks[i];
    return (double)w.depth * (double)w.K * (1.0 + ctx->progs[w.prog].pk.code.size());
  };
  const bool interp_only = pl->cfg.interp_only != 0;
  // walks of more than `slice` primes run as prime slices
  auto kind_of = [&](int i) { return walk_kind(ctx->progs[b.walks[i].prog], interp_only); };
  auto slice_kp = [&](int K) { return (int)pow2_ceil((uint32_t)std::min(K, slice)); };
  // interpreted walks of one order: LDS classes whose LDS size would lower
  // the occupancy of the smaller classes get stages of their own
  std::vector<int> group(n, 0);
  {
    const bool plain_mode = pl->cfg.mode == DRX_MODE_PLAIN;
    for (int m = 2; m <= kMaxM; ++m) {
      size_t cls_lds[4] = {0, 0, 0, 0};
      bool present[4] = {false, false, false, false};
      int kp = 1;
      for (int i = 0; i < n; ++i) {
        const HostProgram& hp = ctx->progs[b.walks[i].prog];
        if ((int)hp.dev.m == m && kind_of(i) == WALK_INTERP) kp = std::max<int>(kp, slice_kp(b.walks[i].K));
      }
      const int thr = plain_mode ? kp : std::max<int>(thr_cfg, kp);
      for (int i = 0; i < n; ++i) {
        const HostProgram& hp = ctx->progs[b.walks[i].prog];
        if ((int)hp.dev.m != m || kind_of(i) != WALK_INTERP) continue;
        const int c = lds_class(hp, WALK_INTERP);
        present[c] = true;
        cls_lds[c] = std::max(cls_lds[c], walk_lds_bytes(hp.dev, WALK_INTERP, m, (uint32_t)b.walks[i].K, thr,
                                                         plain_mode ? 1 : std::min(wpg_cfg, thr / kp), slice));
      }
      int rep[4] = {0, 1, 2, 3};
      int cur = -1;
      size_t cur_lds = 0;
      for (int c = 0; c < 4; ++c) {
        if (!present[c]) continue;
        if (cur >= 0) {
          const size_t merged = std::max(cur_lds, cls_lds[c]);
          const int o0 = walk_blocks_per_cu(m, WALK_INTERP, thr, cur_lds);
          if (o0 > 0 && walk_blocks_per_cu(m, WALK_INTERP, thr, merged) == o0) {
            rep[c] = cur;
            cur_lds = merged;
            continue;
          }
        }
        cur = c;
        cur_lds = cls_lds[c];
      }
      for (int i = 0; i < n; ++i) {
        const HostProgram& hp = ctx->progs[b.walks[i].prog];
        if ((int)hp.dev.m == m && kind_of(i) == WALK_INTERP) group[i] = rep[lds_class(hp, WALK_INTERP)];
      }
    }
  }
  auto gkey = [&](int i) {
    const HostProgram& hp = ctx->progs[b.walks[i].prog];
    return ((int)hp.dev.m * 4 + kind_of(i)) * 4 + group[i];
  };
  auto shape_of = [&](int i) { return walk_shape(ctx->progs[b.walks[i].prog], kind_of(i)); };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    const int kx = gkey(x), ky = gkey(y);
    if (kx != ky) return kx < ky;
    const double cx = cost(x), cy = cost(y);
    if (cx != cy) return cx > cy;
    return shape_of(x) < shape_of(y);
  });

  std::vector<DevWalk> hw(n);
  std::vector<DevTask> ht, hseed;
  std::vector<DevSpan> hs;
  std::vector<DevRange> hr;
  std::vector<double> res0, fst0, axr, axd;
  std::vector<int64_t> axi;
  uint64_t cp_total = 0, mf_total = 0, fdt_total = 0;
  pl->stages.clear();
  pl->walk_order.clear();
  double rsteps = 0, mmt = 0, mat = 0;
  for (int oi = 0; oi < n; ++oi) {
    const int i = order[oi];
    const drx_walk& w = b.walks[i];
    const HostProgram& hp = ctx->progs[w.prog];
    const int m = (int)hp.dev.m, dim = (int)hp.dev.dim, K = w.K;
    const int kind = kind_of(i);
    DevWalk& dw = hw[i];
    dw.prog = w.prog;
    dw.K = K;
    dw.m = m;
    dw.dim = dim;
    dw.target = w.target;
    dw.res_off = res0.size();
    dw.fst_off = fst0.size();
    dw.axr_off = axr.size();
    dw.axi_off = axi.size();
    dw.axd_off = axd.size();
    dw.fdt_off = fdt_total;
    if (kind != WALK_INTERP) fdt_total += (uint64_t)m * m * (fd_dmax(m) + 1) * K;
    // initial state and axes
    const size_t r0 = res0.size();
    res0.resize(r0 + (size_t)m * m * K, 0.0);
    const size_t f0 = fst0.size();
    fst0.resize(f0 + fst_size(m), 0.0);
    int rc0 = init_tables(ctx, w, m, b.init_res_pool, b.init_f_pool, res0.data() + r0, fst0.data() + f0);
    if (rc0) return rc0;
    const size_t a0 = axr.size(), i0 = axi.size(), d0 = axd.size();
    axr.resize(a0 + (size_t)dim * 3 * K);
    axi.resize(i0 + (size_t)dim * 2);
    axd.resize(d0 + (size_t)dim);
    if (dim > 0) {
      rc0 = axis_tables(ctx, b.axis_pool + w.axis_off, dim, K, w.depth, axr.data() + a0, axi.data() + i0,
                        axd.data() + d0);
      if (rc0) return fail(ctx, rc0, "walk %d: %s", i, std::string(ctx->err).c_str());
    }
    // tasks
    const bool plain = pl->cfg.mode == DRX_MODE_PLAIN;
    const int64_t chunk = plain ? std::max<int64_t>(w.depth, 1) : chunk_cfg;
    const int lclass = group[i];
    if (pl->stages.empty() || pl->stages.back().m != m || pl->stages.back().kind != kind ||
        pl->stages.back().lclass != lclass) {
      Stage g;
      g.m = m;
      g.kind = kind;
      g.lclass = lclass;
      g.task_begin = g.task_end = (int)ht.size();
      g.span_begin = g.span_end = (int)hs.size();
      g.seed_begin = g.seed_end = (int)hseed.size();
      g.range_begin = g.range_end = (int)hr.size();
      g.threads = plain ? slice_kp(K) : std::max<int>(thr_cfg, slice_kp(K));
      pl->stages.push_back(g);
    }
    Stage& g = pl->stages.back();
    g.threads = std::max<int>(g.threads, slice_kp(K));
    g.seed_kp = std::max<int>(g.seed_kp, slice_kp(K));
    DevSpan sp{};
    sp.walk = i;
    sp.out_row = i;
    sp.task_begin = (int)ht.size();
    // a companion walk from the identity in one step range, with only p and q
    // asked for: the chunk product is needed in rows 0 and m - 1 only
    const bool rows_only = kind == WALK_FD_COMP && w.init_kind == DRX_INIT_IDENTITY && !b.full &&
                           w.depth <= (int64_t)chunk;
    const int shape = shape_of(i) | (rows_only ? kShapeRows : 0);
    mf_total = (mf_total + 1) & ~(uint64_t)1;  // 16-byte aligned M_f per walk (its tasks stay contiguous)
    for (int64_t s0 = 0; s0 < w.depth; s0 += chunk) {
      DevTask t{};
      t.walk = i;
      t.shape = shape;
      t.s_begin = s0;
      t.s_end = std::min<int64_t>(w.depth, s0 + chunk);
      t.cp_off = cp_total;
      t.mf_off = mf_total;
      cp_total += (uint64_t)m * m * K;
      mf_total += (uint64_t)(t.s_end - t.s_begin) * m * m;
      for (int64_t r = t.s_begin; r < t.s_end; r += kMfRange) {  // M_f once per range: on its first slice
        DevRange rg{};
        rg.task = (int)ht.size();
        rg.s_begin = r;
        rg.s_end = std::min<int64_t>(t.s_end, r + kMfRange);
        hr.push_back(rg);
      }
      for (int k0 = 0; k0 < K; k0 += slice) {
        t.k0 = k0;
        t.kn = std::min(slice, K - k0);
        ht.push_back(t);
        if (kind != WALK_INTERP && s0 == 0) hseed.push_back(t);  // Newton tables: one seed per slice
      }
    }
    sp.task_end = (int)ht.size();
    hs.push_back(sp);
    g.task_end = (int)ht.size();
    g.span_end = (int)hs.size();
    g.seed_end = (int)hseed.size();
    g.range_end = (int)hr.size();
    pl->walk_order.push_back(i);
    rsteps += (double)w.depth * K;
    mmt += (double)w.depth * K * (rows_only ? hp.mm_ops_rows : hp.mm_ops);
    mat += (double)w.depth * K * (rows_only ? hp.ma_ops_rows : hp.ma_ops);
  }
  // LDS per stage (max over its programs), checked against the device limit
  const size_t lds_limit = device_lds_limit(ctx);
  for (auto& g : pl->stages) {
    // the walk kernel's LDS grows with the workgroup: halve it (never below
    // the widest prime stripe) until the stage fits; the sub-chunk split
    // changes, the product does not. Walks per workgroup: as asked, but
    // every lane group at least one prime stripe wide
    for (;;) {
      size_t kp_max = 1;
      for (int si = g.span_begin; si < g.span_end; ++si) kp_max = std::max<size_t>(kp_max, slice_kp(hw[hs[si].walk].K));
      g.wpg = pl->cfg.mode == DRX_MODE_PLAIN ? 1 : std::max(1, std::min(wpg_cfg, g.threads / (int)kp_max));
      g.lds = 0;
      for (int si = g.span_begin; si < g.span_end; ++si) {
        const DevWalk& dw = hw[hs[si].walk];
        g.lds = std::max(g.lds, walk_lds_bytes(ctx->progs[dw.prog].dev, g.kind, g.m, (uint32_t)dw.K, g.threads,
                                               g.wpg, slice));
      }
      if (g.lds <= lds_limit || (size_t)g.threads / 2 < kp_max) break;
      g.threads /= 2;
    }
    for (int si = g.span_begin; si < g.span_end; ++si) {
      const DevWalk& dw = hw[hs[si].walk];
      const DevProgram& dp = ctx->progs[dw.prog].dev;
      g.mf_lds = std::max(g.mf_lds, mf_lds_bytes(dp));
      if (g.kind != WALK_INTERP)
        g.seed_lds = std::max(g.seed_lds, seed_lds_bytes(dp, g.seed_kp, fd_seed_threads(g.m, g.seed_kp)));
    }
    const size_t need = std::max(g.lds, std::max(g.mf_lds, g.seed_lds));
    if (need > lds_limit)
      return fail(ctx, DRX_ERANGE, "program needs %zu bytes of LDS (limit %zu); use fewer threads", need, lds_limit);
  }
  // groups of walks per workgroup: the tasks of a stage with wpg > 1 in plan
  // order (longest first), each appended to the open group of its key -- the
  // interpreter's lane groups must share program and prime slice (one
  // instruction stream and one LDS constant table per workgroup), the
  // forward-difference ones entry kinds, reduced columns, shape and stripe
  // width -- so a group holds tasks of about equal length; groups keep the
  // order in which they were opened
  std::vector<int32_t> hgroup;
  for (Stage& g : pl->stages) {
    if (g.wpg <= 1) continue;
    g.group_begin = (int)(hgroup.size() / g.wpg);
    std::map<std::array<int64_t, 4>, size_t> open;  // key -> first word of its open group
    for (int ti = g.task_begin; ti < g.task_end; ++ti) {
      const DevTask& t = ht[ti];
      const DevProgram& dp = ctx->progs[hw[t.walk].prog].dev;
      const std::array<int64_t, 4> key =
          g.kind == WALK_INTERP
              ? std::array<int64_t, 4>{(int64_t)hw[t.walk].prog, (int64_t)t.k0, (int64_t)t.kn, 0}
              : std::array<int64_t, 4>{(int64_t)dp.kinds, (int64_t)dp.colred, (int64_t)t.shape,
                                       (int64_t)pow2_ceil((uint32_t)t.kn)};
      auto it = open.find(key);
      if (it == open.end()) {
        it = open.emplace(key, hgroup.size()).first;
        hgroup.resize(hgroup.size() + g.wpg, -1);
      }
      size_t slot = it->second;
      while (hgroup[slot] >= 0) ++slot;
      hgroup[slot] = ti;
      if (slot + 1 == it->second + g.wpg) open.erase(it);  // full
    }
    g.group_end = (int)(hgroup.size() / g.wpg);
  }
  if (pl->cfg.mode == DRX_MODE_PERSISTENT) {
    // each stage's walk kernel as a persistent grid of the workgroups the
    // device holds at once, pulling the stage's tasks (longest first) from a
    // queue; the whole sequence is replayed as one captured graph
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || ncu < 1)
      ncu = 1;
    for (Stage& g : pl->stages) {
      // four times the resident workgroups, each retiring after its quota of
      // tasks: a grid of only the resident ones would hold every slot until
      // its queue is empty and starve the FP64 M_f pass on the other stream
      // (mix4: 2.82 s per batch against 1.90 s)
      const int occ = walk_blocks_per_cu(g.m, g.kind, g.threads, g.lds, true);
      g.persist = 4 * std::max(1, occ) * ncu;
    }
    HIPCHK(ctx, pl->d_queue.ensure(std::max<size_t>(1, pl->stages.size()) * 4));
    pl->cfg.use_graph = 1;
  }
  {
    int64_t maxd = 0;
    for (int i = 0; i < n; ++i) maxd = std::max<int64_t>(maxd, b.walks[i].depth);
    shadow_windows(maxd, &pl->win, &pl->nwin);
    order_ranges(pl->stages, hr, pl->win, pl->nwin);
  }
  {
    // the M_f pass first when the walk work dwarfs it: lane instructions of
    // the walk (about 12 per packed residue instruction, 6 per modular
    // multiply of a forward-difference step) against those of the M_f pass
    // (about 12 per FP64 instruction, one lane per step)
    double walk_li = 0, mf_li = 0;
    for (int i = 0; i < n; ++i) {
      const drx_walk& w = b.walks[i];
      const HostProgram& hp = ctx->progs[w.prog];
      const double per = kind_of(i) == WALK_INTERP ? 12.0 * (double)hp.pk.code.size() : 6.0 * hp.mm_ops;
      walk_li += (double)w.depth * w.K * per;
      mf_li += (double)w.depth * 12.0 * (double)hp.pk.fcode.size();
    }
    pl->mf_first = pl->cfg.mode != DRX_MODE_PLAIN && walk_li >= 16.0 * mf_li;
  }
  pl->ntask = (int)ht.size();
  pl->residue_steps = rsteps;
  pl->mm_total = mmt;
  pl->ma_total = mat;

  int rc;
  if ((rc = upload(ctx, pl->d_walks, hw)) || (rc = upload(ctx, pl->d_tasks, ht)) || (rc = upload(ctx, pl->d_seeds, hseed)) ||
      (rc = upload(ctx, pl->d_spans, hs)) ||
      (rc = upload(ctx, pl->d_ranges, hr)) ||
      (rc = upload(ctx, pl->d_res0, res0)) || (rc = upload(ctx, pl->d_fst0, fst0)) || (rc = upload(ctx, pl->d_axr, axr)) ||
      (rc = upload(ctx, pl->d_axi, axi)) || (rc = upload(ctx, pl->d_axd, axd)) ||
      (!hgroup.empty() && (rc = upload(ctx, pl->d_gtask, hgroup))))
    return rc;
  HIPCHK(ctx, pl->d_res.ensure(std::max<size_t>(1, res0.size()) * 8));
  HIPCHK(ctx, pl->d_fst.ensure(std::max<size_t>(1, fst0.size()) * 8));
  HIPCHK(ctx, pl->d_cp.ensure(std::max<uint64_t>(1, cp_total) * 8));
  HIPCHK(ctx, pl->d_mf.ensure((std::max<uint64_t>(16, mf_total) + kMfSlack) * 8));
  HIPCHK(ctx, pl->d_fdt.ensure(std::max<uint64_t>(1, fdt_total) * 8));
  HIPCHK(ctx, pl->d_fout.ensure((size_t)std::max(1, n) * 6 * 8));
  HIPCHK(ctx, pl->d_rout.ensure((size_t)std::max(1, n) * 2 * pl->ld * 8));
  if (b.full) HIPCHK(ctx, pl->d_pout.ensure((size_t)std::max(1, n) * 16 * pl->ld * 8));
  if (b.full_f) HIPCHK(ctx, pl->d_pfull.ensure((size_t)std::max(1, n) * 16 * 8));
  if (b.top_k > 0) {
    if (b.top_k > n) return fail(ctx, DRX_EINVAL, "top_k=%d > n=%d", b.top_k, n);
    HIPCHK(ctx, pl->d_topk.ensure((size_t)b.top_k * 4));
    HIPCHK(ctx, pl->d_topk_ws.ensure(std::max<size_t>(16, topk_workspace_bytes(n, b.top_k))));
  }
  if (b.crt_limbs) {
    std::vector<int32_t> rowk(2 * (size_t)n);
    for (int i = 0; i < n; ++i) rowk[2 * i] = rowk[2 * i + 1] = b.walks[i].K;
    if ((rc = upload(ctx, pl->d_rowk, rowk))) return rc;
    HIPCHK(ctx, pl->d_limbs.ensure((size_t)std::max(1, n) * 2 * pl->ld * 4));
    HIPCHK(ctx, pl->d_neg.ensure((size_t)std::max(1, n) * 2));
  }
  if ((rc = upload_programs(ctx))) return rc;

  pl->early_out = sinks_page_locked(b);
  KArgs& a = pl->args;
  a.tasks = pl->d_tasks.as<DevTask>();
  a.queue = pl->d_queue.p ? pl->d_queue.as<uint32_t>() : nullptr;
  a.gtask = hgroup.empty() ? nullptr : pl->d_gtask.as<int32_t>();
  a.wpg = 1;
  a.walks = pl->d_walks.as<DevWalk>();
  a.progs = ctx->d_progs.as<DevProgram>();
  a.spans = pl->d_spans.as<DevSpan>();
  a.ranges = pl->d_ranges.as<DevRange>();
  a.invs = ctx->d_invs.as<double>();
  a.fdt = pl->d_fdt.as<double>();
  a.code = ctx->d_code.as<uint32_t>();
  a.cres = ctx->d_cres.as<double>();
  a.cflt = ctx->d_cflt.as<double>();
  a.prime_p = ctx->d_prime_p.as<double>();
  a.prime_pinv = ctx->d_prime_pinv.as<double>();
  a.prime_e = ctx->d_prime_e.as<uint32_t>();
  a.axr = pl->d_axr.as<double>();
  a.axi = pl->d_axi.as<int64_t>();
  a.axd = pl->d_axd.as<double>();
  a.res_in = pl->d_res0.as<double>();
  a.fst_in = pl->d_fst0.as<double>();
  a.res = pl->d_res.as<double>();
  a.fst = pl->d_fst.as<double>();
  a.cp = pl->d_cp.as<double>();
  a.mf = pl->d_mf.as<double>();
  a.fout = pl->d_fout.as<double>();
  a.rout = pl->d_rout.as<int64_t>();
  a.pout = b.full ? pl->d_pout.as<int64_t>() : nullptr;
  a.fcompact = nullptr;
  a.ffull = b.full_f ? pl->d_pfull.as<double>() : nullptr;
  a.out_ld = (uint64_t)pl->ld;
  GarnerArgs& g = pl->gargs;
  g.res = pl->d_rout.as<int64_t>();
  g.res_stride = (uint64_t)pl->ld;
  g.n = b.crt_limbs ? 2 * n : 0;
  g.K = kmax;
  g.row_k = b.crt_limbs ? pl->d_rowk.as<int32_t>() : nullptr;
  g.prime_p = ctx->d_prime_p.as<double>();
  g.prime_pinv = ctx->d_prime_pinv.as<double>();
  g.prime_u = ctx->d_prime_u.as<uint32_t>();
  g.inv_table = ctx->d_inv.as<uint32_t>();
  g.inv_ld = (uint64_t)ctx->kinv;
  g.limbs = pl->d_limbs.as<uint32_t>();
  g.neg = pl->d_neg.as<uint8_t>();
  if ((rc = setup_crt(ctx, pl))) return rc;

  if (pl->cfg.use_graph && (rc = capture_plan(pl))) return rc;
  return DRX_OK;
}

int plan_fetch(drx_plan* pl) {
  drx_ctx* ctx = pl->ctx;
  const drx_batch& b = pl->b;
  hipStream_t s = ctx->stream;
  const int n = pl->n;
  if (b.out) HIPCHK(ctx, hipMemcpyAsync(b.out, pl->d_fout.p, (size_t)n * 6 * 8, hipMemcpyDeviceToHost, s));
  if (b.residues)
    HIPCHK(ctx, hipMemcpyAsync(b.residues, pl->d_rout.p, (size_t)n * 2 * pl->ld * 8, hipMemcpyDeviceToHost, s));
  if (b.full) HIPCHK(ctx, hipMemcpyAsync(b.full, pl->d_pout.p, (size_t)n * 16 * pl->ld * 8, hipMemcpyDeviceToHost, s));
  if (b.full_f) HIPCHK(ctx, hipMemcpyAsync(b.full_f, pl->d_pfull.p, (size_t)n * 16 * 8, hipMemcpyDeviceToHost, s));
  if (b.crt_limbs) {
    HIPCHK(ctx, hipMemcpyAsync(b.crt_limbs, pl->d_limbs.p, (size_t)n * 2 * pl->ld * 4, hipMemcpyDeviceToHost, s));
    if (b.crt_neg) HIPCHK(ctx, hipMemcpyAsync(b.crt_neg, pl->d_neg.p, (size_t)n * 2, hipMemcpyDeviceToHost, s));
  }
  if (b.top_k > 0 && b.topk_idx)
    HIPCHK(ctx, hipMemcpyAsync(b.topk_idx, pl->d_topk.p, (size_t)b.top_k * 4, hipMemcpyDeviceToHost, s));
  return DRX_OK;
}

int plan_run_one(drx_plan* pl, bool timed, bool fetch = true) {
  drx_ctx* ctx = pl->ctx;
  if (pl->d_pout.p && pl->b.full)
    HIPCHK(ctx, hipMemsetAsync(pl->d_pout.p, 0, (size_t)pl->n * 16 * pl->ld * 8, ctx->stream));
  // A run that copies into page-locked sinks launches on streams, its copies
  // beside the shadow chain: a graph with the copies captured as nodes
  // measured 0.95 ms per pcf2 step, the replay followed by the copies 0.71-0.73,
  // the stream launch 0.68-0.70.
  if (pl->gexec && !(fetch && pl->early_out && !pl->vec)) {
    if (timed) HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    HIPCHK(ctx, hipGraphLaunch(pl->gexec, ctx->stream));
    if (timed) HIPCHK(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
    return fetch ? plan_fetch(pl) : DRX_OK;
  }
  return launch_sequence(pl, timed, fetch);
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

int drx_abi_version(void) { return DRX_ABI_VERSION; }

int drx_device_count(int* n) {
  if (!n) return DRX_EINVAL;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return DRX_OK;
}

int drx_ctx_create(int device, drx_ctx** out) {
  if (!out) return DRX_EINVAL;
  *out = nullptr;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) return DRX_EHIP;
  if (device < 0 || device >= nd) return DRX_EINVAL;
  drx_ctx* ctx = new (std::nothrow) drx_ctx();
  if (!ctx) return DRX_ENOMEM;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream3, hipStreamNonBlocking) != hipSuccess) {
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    delete ctx;
    return DRX_EHIP;
  }
  for (auto& e : ctx->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      drx_ctx_destroy(ctx);
      return DRX_EHIP;
    }
  }
  if (noalloc_init(ctx) != DRX_OK) {
    drx_ctx_destroy(ctx);
    return DRX_EHIP;
  }
  *out = ctx;
  return DRX_OK;
}

void drx_ctx_destroy(drx_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
  if (ctx->stream3) (void)hipStreamSynchronize(ctx->stream3);
  for (DBuf* d : {&ctx->d_prime_p, &ctx->d_prime_pinv, &ctx->d_prime_e, &ctx->d_prime_u, &ctx->d_invs, &ctx->d_inv, &ctx->d_code,
                  &ctx->d_cres, &ctx->d_cflt, &ctx->d_progs, &ctx->d_tmp0, &ctx->d_tmp1, &ctx->d_tmp2, &ctx->d_tmp3,
                  &ctx->d_crt_rows, &ctx->d_crt_a, &ctx->d_crt_c, &ctx->d_crt_k})
    d->release();
  retire_crt_tables(ctx);
  for (void* p : ctx->retired) (void)hipFree(p);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
  if (ctx->stream3) (void)hipStreamDestroy(ctx->stream3);
  for (hipStream_t& m : ctx->mstream)
    if (m) (void)hipStreamDestroy(m);
  delete ctx;
}

const char* drx_last_error(const drx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int drx_set_primes(drx_ctx* ctx, const int64_t* primes, int K) {
  if (!ctx) return DRX_EINVAL;
  if (!primes || K < 1 || K > kMaxK) return fail(ctx, DRX_EINVAL, "K=%d outside [1, %d]", K, kMaxK);
  std::vector<uint32_t> pr(K);
  int kwalk = 0;
  {
    std::vector<uint32_t> sorted;
    sorted.reserve(K);
    for (int i = 0; i < K; ++i) {
      const int64_t p = primes[i];
      if (p < 3 || p >= ((int64_t)1 << 31) || !is_prime_u32((uint32_t)p))
        return fail(ctx, DRX_EINVAL, "prime %d (%lld) is not an odd prime below 2^31", i, (long long)p);
      pr[i] = (uint32_t)p;
      if (kwalk == i && p > ((int64_t)1 << 30)) ++kwalk;
      sorted.push_back((uint32_t)p);
    }
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
      return fail(ctx, DRX_EINVAL, "a prime is repeated");
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (!ctx->primes.empty() && (size_t)K >= ctx->primes.size() &&
      std::equal(ctx->primes.begin(), ctx->primes.end(), pr.begin()) && (size_t)K == ctx->primes.size())
    return DRX_OK;
  if (!ctx->progs.empty()) {
    // programs hold residues for the old table; only an extension is allowed
    if ((size_t)K < ctx->primes.size() || !std::equal(ctx->primes.begin(), ctx->primes.end(), pr.begin()))
      return fail(ctx, DRX_ESTATE, "primes changed while programs are loaded; release programs first");
  }
  std::vector<double> pd(K), pinv(K);
  std::vector<uint32_t> pe(K);
  for (int i = 0; i < K; ++i) {
    pd[i] = (double)pr[i];
    pinv[i] = 1.0 / (double)pr[i];
    pe[i] = pr[i] - 2;
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  // an extension keeps every earlier plan valid; any other change retires them
  const bool extends = pr.size() >= ctx->primes.size() && std::equal(ctx->primes.begin(), ctx->primes.end(), pr.begin());
  if (!extends) {
    ++ctx->epoch;
    retire_crt_tables(ctx);
    ctx->garner_w.clear();  // its device copy is rewritten (retired if outgrown) on next use
  }
  HIPCHK(ctx, ctx->d_prime_p.ensure_keep(K * 8, ctx->retired));
  HIPCHK(ctx, ctx->d_prime_pinv.ensure_keep(K * 8, ctx->retired));
  HIPCHK(ctx, ctx->d_prime_e.ensure_keep(K * 4, ctx->retired));
  HIPCHK(ctx, ctx->d_prime_u.ensure_keep(K * 4, ctx->retired));
  HIPCHK(ctx, hipMemcpy(ctx->d_prime_p.p, pd.data(), K * 8, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(ctx->d_prime_pinv.p, pinv.data(), K * 8, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(ctx->d_prime_e.p, pe.data(), K * 4, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(ctx->d_prime_u.p, pr.data(), K * 4, hipMemcpyHostToDevice));
  {
    std::vector<double> invs((size_t)K * (kFdInv + 1), 0.0);
    for (int i = 0; i < K; ++i)
      for (int t = 1; t <= kFdInv; ++t) invs[(size_t)i * (kFdInv + 1) + t] = balanced(inv_mod((uint32_t)t, pr[i]), pr[i]);
    HIPCHK(ctx, ctx->d_invs.ensure_keep(invs.size() * 8, ctx->retired));
    HIPCHK(ctx, hipMemcpy(ctx->d_invs.p, invs.data(), invs.size() * 8, hipMemcpyHostToDevice));
  }
  ctx->primes.swap(pr);
  ctx->kinv = 0;
  ctx->kwalk = kwalk;
  // both Garner tables now, so that no CRT call (drx_crt_noalloc above all)
  // has to build or allocate them
  int rc = ensure_inv_table(ctx, std::min(K, kCrtMatrixMinK - 1));
  if (!rc) rc = ensure_garner_w(ctx, K);
  return rc;
}

int drx_program_upload(drx_ctx* ctx, const uint32_t* ops, int n_ops, int m, int dim, const int64_t* consts,
                       const uint32_t* const_res, const double* const_f, int n_const, int kcap, int* prog_id) {
  return drx_program_upload_mode(ctx, ops, n_ops, m, dim, consts, const_res, const_f, n_const, kcap,
                                 DRX_WALK_MATRIX, prog_id);
}

int drx_program_upload_mode(drx_ctx* ctx, const uint32_t* ops, int n_ops, int m, int dim, const int64_t* consts,
                            const uint32_t* const_res, const double* const_f, int n_const, int kcap, int mode,
                            int* prog_id) {
  if (!ctx) return DRX_EINVAL;
  if (mode != DRX_WALK_MATRIX && mode != DRX_WALK_VECTOR) return fail(ctx, DRX_EINVAL, "unknown walk mode %d", mode);
  if (!prog_id || (n_ops > 0 && !ops) || n_ops < 0 || n_const < 0)
    return fail(ctx, DRX_EINVAL, "bad program arguments");
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (ctx->primes.empty()) return fail(ctx, DRX_ESTATE, "set primes before uploading programs");
  if (n_const > 0 && !consts && !(const_res && const_f))
    return fail(ctx, DRX_EINVAL, "constants missing (give consts or const_res + const_f)");
  const int K = (int)ctx->primes.size();
  const bool given = const_res != nullptr;
  // constant residues for the first kcap primes (all primes set when kcap <= 0 and
  // the constants are given as int64): walks of this program may use up to kc primes
  const int kc = (given || kcap > 0) ? std::min(kcap, K) : K;
  if (given && kcap < 1) return fail(ctx, DRX_EINVAL, "kcap must be >= 1");
  std::vector<char> is_one(n_const, 0), is_zero(n_const, 0);
  for (int i = 0; i < n_const; ++i) {
    if (given) {
      bool one = const_f[i] == 1.0, zero = const_f[i] == 0.0;
      for (int k = 0; k < kc; ++k) {
        one = one && const_res[(size_t)i * kcap + k] == 1u;
        zero = zero && const_res[(size_t)i * kcap + k] == 0u;
      }
      is_one[i] = one;
      is_zero[i] = zero;
    } else {
      is_one[i] = consts[i] == 1;
      is_zero[i] = consts[i] == 0;
    }
  }
  HostProgram hp;
  hp.mode = mode;
  const std::string err = pack_program(ops, n_ops, m, dim, n_const, is_one, is_zero, &hp.pk,
                                       mode == DRX_WALK_VECTOR ? kPackVector : kPackMatrix);
  if (!err.empty()) return fail(ctx, DRX_EINVAL, "%s", err.c_str());
  const PackedProgram& pk = hp.pk;
  const int nc = (int)pk.const_slots.size();
  // every constant the code may reference: the given ones, then the folded
  // ones computed from them with the operation the code would have applied
  // (residues mod each covered prime, FP64 with the same IEEE operation)
  const int n_all = n_const + (int)pk.derived.size();
  std::vector<uint32_t> cr((size_t)n_all * kc);
  std::vector<double> cf(n_all);
  for (int ci = 0; ci < n_const; ++ci) {
    for (int k = 0; k < kc; ++k) {
      const uint32_t p = ctx->primes[k];
      const uint32_t r = given ? const_res[(size_t)ci * kcap + k] : mod_i64(consts[ci], p);
      if (r >= p) return fail(ctx, DRX_EINVAL, "constant %d: residue not canonical", ci);
      cr[(size_t)ci * kc + k] = r;
    }
    cf[ci] = given ? const_f[ci] : (double)consts[ci];
  }
  for (size_t j = 0; j < pk.derived.size(); ++j) {
    const DerivedConst& dc = pk.derived[j];
    const int ci = n_const + (int)j;
    const double fa = cf[dc.a], fb = dc.b >= 0 ? cf[dc.b] : 0.0;
    switch (dc.op) {
      case D_ADD: cf[ci] = fa + fb; break;
      case D_SUB: cf[ci] = fa - fb; break;
      case D_MUL: cf[ci] = fa * fb; break;
      case D_NEG: cf[ci] = -fa; break;
      case D_SQR: cf[ci] = fa * fa; break;
      default: cf[ci] = std::fabs(fa) > 1e-300 ? 1.0 / fa : 0.0; break;  // D_INV
    }
    for (int k = 0; k < kc; ++k) {
      const uint64_t p = ctx->primes[k];
      const uint64_t ra = cr[(size_t)dc.a * kc + k], rb = dc.b >= 0 ? cr[(size_t)dc.b * kc + k] : 0;
      uint64_t r;
      switch (dc.op) {
        case D_ADD: r = (ra + rb) % p; break;
        case D_SUB: r = (ra + p - rb) % p; break;
        case D_MUL: r = ra * rb % p; break;
        case D_NEG: r = (p - ra) % p; break;
        case D_SQR: r = ra * ra % p; break;
        default: r = inv_mod((uint32_t)ra, (uint32_t)p); break;  // D_INV, 0 -> 0
      }
      cr[(size_t)ci * kc + k] = (uint32_t)r;
    }
  }
  DevProgram& d = hp.dev;
  // the residue code and then the FP64 code in the pool, each from a 64-byte
  // boundary, kResWords words per instruction (the fields unpacked, so the
  // interpreters use them with no decoding)
  while (ctx->h_code.size() % 16) ctx->h_code.push_back(0);
  d.code_off = (uint32_t)ctx->h_code.size();
  d.ncode = (uint32_t)pk.code.size();
  d.fcode_off = (uint32_t)(d.code_off + ((kResWords * pk.code.size() + 15) & ~(size_t)15));
  d.nfcode = (uint32_t)pk.fcode.size();
  d.nconst = (uint32_t)nc;
  d.kcap = (uint32_t)kc;
  d.cres_off = ctx->h_cres.size();
  d.cflt_off = ctx->h_cflt.size();
  d.m = (uint32_t)m;
  d.dim = (uint32_t)dim;
  d.kinds = pk.kinds;
  d.colred = pk.colred;
  d.const_lds = (uint32_t)pk.const_lds;
  d.nslots = (uint32_t)pk.nslots;
  d.fnslots = (uint32_t)pk.fnslots;
  d.pad = 0;
  d.fd_ok = (uint32_t)pk.fd_ok;
  d.fd_degree = (uint32_t)pk.fd_degree;
  d.fd_deg = pk.fd_deg;
  hp.fd = mode == DRX_WALK_MATRIX && pk.fd_ok &&
          ((m == 2 && pk.fd_degree <= kFdMax2) || (m == 3 && pk.fd_degree <= kFdMax3) ||
           (m == 4 && comp_shape(d) >= 0));
  for (const uint32_t w : pk.code) {
    const ResWords u = unpack_res(w);
    ctx->h_code.insert(ctx->h_code.end(), u.w, u.w + kResWords);
  }
  while (ctx->h_code.size() % 16) ctx->h_code.push_back(0);
  for (const uint32_t w : pk.fcode) {
    const ResWords u = unpack_res(w);
    ctx->h_code.insert(ctx->h_code.end(), u.w, u.w + kResWords);
  }
  for (int s = 0; s < nc; ++s) {
    const int ci = pk.const_slots[s];
    for (int k = 0; k < kc; ++k) ctx->h_cres.push_back(balanced(cr[(size_t)ci * kc + k], ctx->primes[k]));
    ctx->h_cflt.push_back(cf[ci]);
  }
  // executed modular work per step: program + structured P*M
  int mmo = 0, mao = 0;
  for (const uint32_t w : pk.code) {
    const uint32_t op = w & 15u;
    if (op == RO_MUL) mmo += 1;
    else if (op == RO_INV) mmo += 60;
    else if (op == RO_NZ1 || op == RO_ZSEL || op == RO_RED) mmo += 1;  // one reduction and a select
    else if (op == RO_MULNZ) mmo += 2;
    else if (op == RO_ADD || op == RO_LIN) mao += 1;
  }
  // the step product (a vector walk's product is part of its code)
  for (int j = 0; j < m && mode == DRX_WALK_MATRIX; ++j) {
    int terms = 0;
    for (int k = 0; k < m; ++k) {
      const uint32_t kd = (pk.kinds >> (2 * (k * m + j))) & 3u;
      if (kd == E_GEN) mmo += m;
      if (kd != E_ZERO) ++terms;
    }
    if (terms > 1) mao += m * (terms - 1);
  }
  if (hp.fd) {
    // per step: the entry tables advance (one add per difference), the
    // interpreted program runs only when the tables are seeded
    mmo = 0;
    mao = 0;
    int tab = 0, mm1 = 0, ma1 = 0;  // table additions; product work per carried row
    for (int e = 0; e < m * m; ++e)
      if (((pk.kinds >> (2 * e)) & 3u) == E_GEN) tab += (int)((pk.fd_deg >> (4 * e)) & 15u);
    for (int j = 0; j < m; ++j) {
      int terms = 0;
      for (int k = 0; k < m; ++k) {
        const uint32_t kd = (pk.kinds >> (2 * (k * m + j))) & 3u;
        if (kd == E_GEN) mm1 += 1;
        if (kd != E_ZERO) ++terms;
      }
      if (terms > 1) ma1 += terms - 1;
    }
    mmo = m * mm1;
    mao = tab + m * ma1;
    hp.mm_ops_rows = 2 * mm1;
    hp.ma_ops_rows = tab + 2 * ma1;
  }
  hp.mm_ops = mmo;
  hp.ma_ops = mao;
  if (!hp.fd) {
    hp.mm_ops_rows = mmo;
    hp.ma_ops_rows = mao;
  }
  ctx->progs.push_back(std::move(hp));
  ctx->progs_dirty = true;
  *prog_id = (int)ctx->progs.size() - 1;
  return DRX_OK;
}

int drx_program_info(drx_ctx* ctx, int prog_id, int* n_dev_ops, int* n_const_used, uint32_t* entry_kinds) {
  if (!ctx) return DRX_EINVAL;
  if (prog_id < 0 || prog_id >= (int)ctx->progs.size()) return fail(ctx, DRX_EINVAL, "bad program id");
  const HostProgram& hp = ctx->progs[prog_id];
  if (n_dev_ops) *n_dev_ops = (int)hp.dev.ncode;
  if (n_const_used) *n_const_used = (int)hp.dev.nconst;
  if (entry_kinds) *entry_kinds = hp.dev.kinds;
  return DRX_OK;
}

int drx_program_pack(const uint32_t* ops, int n_ops, int m, int dim, const int64_t* consts, int n_const, int mode,
                     uint32_t* words, int max_words, int* const_slots, int max_slots, int32_t* derived,
                     int max_derived, int64_t* meta, int max_meta, char* err, int err_len) {
  if ((n_ops > 0 && !ops) || n_ops < 0 || n_const < 0 || (n_const > 0 && !consts) || !meta || max_meta < 0)
    return DRX_EINVAL;
  std::vector<char> is_one(n_const, 0), is_zero(n_const, 0);
  for (int i = 0; i < n_const; ++i) {
    is_one[i] = consts[i] == 1;
    is_zero[i] = consts[i] == 0;
  }
  PackedProgram pk;
  if (mode != DRX_WALK_MATRIX && mode != DRX_WALK_VECTOR) return DRX_EINVAL;
  const std::string msg = pack_program(ops, n_ops, m, dim, n_const, is_one, is_zero, &pk,
                                       mode == DRX_WALK_VECTOR ? kPackVector : kPackMatrix);
  if (!msg.empty()) {
    if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", msg.c_str());
    return DRX_EINVAL;
  }
  const int64_t mv[DRX_PACK_META] = {(int64_t)pk.code.size(), pk.kinds, pk.colred, pk.const_lds, pk.nslots,
                                 (int64_t)pk.const_slots.size(), pk.fd_ok, pk.fd_degree,
                                 (int64_t)pk.fcode.size(), pk.fnslots, (int64_t)pk.derived.size(),
                                 pk.n_inv_batched, pk.n_src_ops};
  for (int i = 0; i < std::min(max_meta, DRX_PACK_META); ++i) meta[i] = mv[i];
  if (words) {
    std::vector<uint32_t> all(pk.code);
    if (all.size() % 2) all.push_back(0);  // the FP64 code starts on an even word, as in the code pool
    all.insert(all.end(), pk.fcode.begin(), pk.fcode.end());
    std::copy(all.begin(), all.begin() + std::min<size_t>(all.size(), std::max(0, max_words)), words);
  }
  if (const_slots)
    for (size_t i = 0; i < pk.const_slots.size() && (int)i < max_slots; ++i) const_slots[i] = pk.const_slots[i];
  if (derived)
    for (size_t i = 0; i < pk.derived.size() && (int)i < max_derived; ++i) {
      derived[3 * i] = (int32_t)pk.derived[i].op;
      derived[3 * i + 1] = pk.derived[i].a;
      derived[3 * i + 2] = pk.derived[i].b;
    }
  return DRX_OK;
}

int drx_program_release_all(drx_ctx* ctx) {
  if (!ctx) return DRX_EINVAL;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->progs.clear();
  ctx->h_code.clear();
  ctx->h_cres.clear();
  ctx->h_cflt.clear();
  ctx->progs_dirty = true;
  ++ctx->epoch;
  return DRX_OK;
}

int drx_plan_create(drx_ctx* ctx, const drx_batch* b, const drx_launch_cfg* cfg, drx_plan** out) {
  if (!ctx) return DRX_EINVAL;
  if (!b || !out || b->n < 0 || (b->n > 0 && (!b->walks || !b->out || !b->residues)))
    return fail(ctx, DRX_EINVAL, "bad batch arguments");
  *out = nullptr;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  drx_plan* pl = new (std::nothrow) drx_plan();
  if (!pl) return fail(ctx, DRX_ENOMEM, "out of host memory");
  pl->ctx = ctx;
  pl->epoch = ctx->epoch;
  pl->b = *b;
  if (cfg) pl->cfg = *cfg;
  if (pl->cfg.chain_cus < 0 || (pl->cfg.chain_cus > 0 && (pl->cfg.use_graph || pl->cfg.mode != DRX_MODE_BATCHED))) {
    delete pl;
    return fail(ctx, DRX_EINVAL, "chain_cus takes the batched mode without graph replay");
  }
  if (pl->cfg.mode == DRX_MODE_PLAIN) {
    // one sub-plan per walk: each run launches and synchronises one walk
    pl->n = b->n;
    pl->ld = b->res_stride;
    for (int i = 0; i < b->n; ++i) {
      drx_batch sb = *b;
      sb.walks = b->walks + i;
      sb.n = 1;
      sb.out = b->out + i;
      sb.residues = b->residues + (size_t)i * 2 * b->res_stride;
      sb.full = b->full ? b->full + (size_t)i * 16 * b->res_stride : nullptr;
      sb.full_f = b->full_f ? b->full_f + (size_t)i * 16 : nullptr;
      sb.crt_limbs = b->crt_limbs ? b->crt_limbs + (size_t)i * 2 * b->res_stride : nullptr;
      sb.crt_neg = b->crt_neg ? b->crt_neg + 2 * i : nullptr;
      sb.top_k = 0;
      sb.topk_idx = nullptr;
      drx_plan* sp = new (std::nothrow) drx_plan();
      if (!sp) {
        delete pl;
        return fail(ctx, DRX_ENOMEM, "out of host memory");
      }
      sp->ctx = ctx;
      sp->b = sb;
      sp->cfg = pl->cfg;
      sp->cfg.use_graph = 0;
      pl->plain.push_back(sp);
      const int rc = plan_build(ctx, sp);
      if (rc) {
        delete pl;
        return rc;
      }
      pl->residue_steps += sp->residue_steps;
      pl->mm_total += sp->mm_total;
      pl->ma_total += sp->ma_total;
      pl->ntask += sp->ntask;
    }
    if (b->top_k > 0) {
      if (b->top_k > b->n) {
        delete pl;
        return fail(ctx, DRX_EINVAL, "top_k=%d > n=%d", b->top_k, b->n);
      }
      if (ctx->d_tmp0.ensure((size_t)b->n * 8) != hipSuccess || ctx->d_tmp1.ensure((size_t)b->top_k * 4) != hipSuccess) {
        delete pl;
        return fail(ctx, DRX_ENOMEM, "device allocation failed");
      }
    }
    *out = pl;
    return DRX_OK;
  }
  const int rc = plan_build(ctx, pl);
  if (rc) {
    delete pl;
    return rc;
  }
  *out = pl;
  return DRX_OK;
}

namespace {
// CU-masked copies of the context's three streams for a split of `n` CUs:
// the FP64 streams (M_f pass, shadow) on n CUs chosen so that every group of
// 32 consecutive CU ids and every residue of the id mod 8 holds the same
// number of them (the chip's XCDs get equal shares whichever of the two
// numberings the ids follow), the residue stream on the rest.
int ensure_masked_streams(drx_ctx* ctx, int n) {
  if (ctx->masked_cus == n && ctx->mstream[0]) return DRX_OK;
  int ncu = 0;
  HIPCHK(ctx, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
  if (ncu < 64 || ncu % 32 != 0 || n < 1 || n >= ncu || ncu % n != 0 || (ncu / n) > 8 || 8 % (ncu / n) != 0)
    return fail(ctx, DRX_EINVAL, "chain_cus=%d: takes a divisor of the %d CUs that leaves 1/2 .. 1/8 of them", n, ncu);
  const int step = ncu / n;
  std::vector<uint32_t> mask_fp((ncu + 31) / 32, 0u), mask_res((ncu + 31) / 32, 0u);
  for (int i = 0; i < ncu; ++i) {
    const bool fp = ((i % 8) + (i / 32)) % step == 0;
    (fp ? mask_fp : mask_res)[i / 32] |= 1u << (i % 32);
  }
  for (hipStream_t& m : ctx->mstream)
    if (m) {
      HIPCHK(ctx, hipStreamSynchronize(m));
      HIPCHK(ctx, hipStreamDestroy(m));
      m = nullptr;
    }
  ctx->masked_cus = 0;
  HIPCHK(ctx, hipExtStreamCreateWithCUMask(&ctx->mstream[0], (uint32_t)mask_res.size(), mask_res.data()));
  HIPCHK(ctx, hipExtStreamCreateWithCUMask(&ctx->mstream[1], (uint32_t)mask_fp.size(), mask_fp.data()));
  HIPCHK(ctx, hipExtStreamCreateWithCUMask(&ctx->mstream[2], (uint32_t)mask_fp.size(), mask_fp.data()));
  ctx->masked_cus = n;
  return DRX_OK;
}

// The context's streams for the duration of one run (restored on scope exit).
struct StreamSwap {
  drx_ctx* ctx;
  hipStream_t saved[3];
  explicit StreamSwap(drx_ctx* c) : ctx(c), saved{c->stream, c->stream2, c->stream3} {
    ctx->stream = ctx->mstream[0];
    ctx->stream2 = ctx->mstream[1];
    ctx->stream3 = ctx->mstream[2];
  }
  ~StreamSwap() {
    ctx->stream = saved[0];
    ctx->stream2 = saved[1];
    ctx->stream3 = saved[2];
  }
};

int plan_run(drx_plan* pl, bool fetch) {
  if (!pl) return DRX_EINVAL;
  drx_ctx* ctx = pl->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (pl->epoch != ctx->epoch)
    return fail(ctx, DRX_ESTATE, "the context's primes or programs changed since this plan was built; build it again");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<StreamSwap> swap;
  if (pl->cfg.chain_cus > 0 && !pl->gexec && pl->plain.empty() && pl->cfg.mode != DRX_MODE_PLAIN) {
    const int rc = ensure_masked_streams(ctx, pl->cfg.chain_cus);
    if (rc) return rc;
    swap.reset(new StreamSwap(ctx));
  }
  Timing tm;
  if (!pl->plain.empty() || pl->cfg.mode == DRX_MODE_PLAIN) {
    const double t0 = now_s();
    for (drx_plan* sp : pl->plain) {
      int rc = plan_run_one(sp, true);
      if (rc) return rc;
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      float a = 0, b2 = 0, c = 0, d = 0;
      (void)hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]);
      (void)hipEventElapsedTime(&b2, ctx->ev[1], ctx->ev[2]);
      (void)hipEventElapsedTime(&c, ctx->ev[2], ctx->ev[3]);
      (void)hipEventElapsedTime(&d, ctx->ev[3], ctx->ev[4]);
      tm.walk_ms += a;
      tm.combine_ms += b2;
      tm.shadow_ms += c;
      tm.tail_ms += d;
      tm.walk_launches += (int)sp->stages.size();
    }
    if (pl->b.top_k > 0) {
      // rank the per-walk deltas on the device like the batched path does
      std::vector<double> dl(pl->n);
      for (int i = 0; i < pl->n; ++i) dl[i] = pl->b.out[i].delta_float;
      HIPCHK(ctx, hipMemcpyAsync(ctx->d_tmp0.p, dl.data(), dl.size() * 8, hipMemcpyHostToDevice, ctx->stream));
      HIPCHK(ctx, ctx->d_tmp2.ensure(std::max<size_t>(16, topk_workspace_bytes(pl->n, pl->b.top_k))));
      HIPCHK(ctx, launch_topk(ctx->d_tmp0.as<double>(), 1, pl->n, pl->b.top_k, ctx->d_tmp1.as<int32_t>(),
                              ctx->d_tmp2.p, ctx->stream));
      HIPCHK(ctx, hipMemcpyAsync(pl->b.topk_idx, ctx->d_tmp1.p, (size_t)pl->b.top_k * 4, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    tm.total_ms = (now_s() - t0) * 1e3;
    ctx->timing = tm;
    return DRX_OK;
  }
  int rc = plan_run_one(pl, true, fetch);
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  float a = 0, b2 = 0, c = 0, d = 0, tot = 0;
  if (!pl->gexec) {
    (void)hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]);
    (void)hipEventElapsedTime(&b2, ctx->ev[1], ctx->ev[2]);
    (void)hipEventElapsedTime(&c, ctx->ev[2], ctx->ev[3]);
    (void)hipEventElapsedTime(&d, ctx->ev[3], ctx->ev[4]);
  }
  (void)hipEventElapsedTime(&tot, ctx->ev[0], ctx->ev[4]);
  tm.walk_ms = a;
  tm.combine_ms = b2;
  tm.shadow_ms = c;
  tm.tail_ms = d;
  tm.total_ms = tot;
  tm.walk_launches = (int)pl->stages.size();
  ctx->timing = tm;
  return DRX_OK;
}
}  // namespace

int drx_plan_run(drx_plan* pl) { return plan_run(pl, true); }

int drx_plan_run_resident(drx_plan* pl) { return plan_run(pl, false); }

int drx_plan_fetch(drx_plan* pl) {
  if (!pl) return DRX_EINVAL;
  drx_ctx* ctx = pl->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (!pl->plain.empty() || pl->cfg.mode == DRX_MODE_PLAIN) return DRX_OK;  // plain runs copy back per walk
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int rc = plan_fetch(pl);
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DRX_OK;
}

namespace {

// The walks of `b` at indices 0, stride, 2 stride, ... as a batch writing
// into host sinks owned by the caller (residues and the FP64 outputs; no CRT
// or top-k).
struct TuneSample {
  std::vector<drx_walk> walks;
  std::vector<drx_walk_out> out;
  std::vector<int64_t> res;
  drx_batch batch{};
  TuneSample(const drx_batch* b, int want) {
    const int stride = (b->n + want - 1) / want;
    int kmax = 1;
    for (int i = 0; i < b->n; i += stride) {
      walks.push_back(b->walks[i]);
      kmax = std::max(kmax, b->walks[i].K);
    }
    out.resize(walks.size());
    res.resize(walks.size() * 2 * (size_t)kmax);
    batch.walks = walks.data();
    batch.n = (int)walks.size();
    batch.axis_pool = b->axis_pool;
    batch.init_res_pool = b->init_res_pool;
    batch.init_f_pool = b->init_f_pool;
    batch.out = out.data();
    batch.residues = res.data();
    batch.res_stride = kmax;
  }
};

// Wall time of one plan run (seconds), best of `reps` after a warm-up.
int time_plan(drx_ctx* ctx, const drx_batch* b, const drx_launch_cfg* c, int reps, double* sec) {
  drx_plan* pl = nullptr;
  int rc = drx_plan_create(ctx, b, c, &pl);
  if (rc) return rc;
  double best = 1e300;
  for (int r = 0; r <= reps && rc == 0; ++r) {
    const double t0 = now_s();
    rc = drx_plan_run(pl);
    if (rc == 0 && r > 0) best = std::min(best, now_s() - t0);
  }
  drx_plan_destroy(pl);
  *sec = best;
  return rc;
}

}  // namespace

namespace {
// The narrowest walk of a sample (its prime stripe bounds walks per workgroup).
int kmin_walk(const std::vector<drx_walk>& walks) {
  int k = INT32_MAX;
  for (const drx_walk& w : walks) k = std::min(k, w.K);
  return walks.empty() ? 1 : k;
}
}  // namespace

int drx_autotune(drx_ctx* ctx, const drx_batch* b, const drx_launch_cfg* base, int max_walks, double budget_s,
                 drx_launch_cfg* best, double* best_ms) {
  if (!ctx || !b || !best || !best_ms || b->n < 1 || !b->walks) return DRX_EINVAL;
  drx_launch_cfg cfg0{};
  if (base) cfg0 = *base;
  cfg0.mode = DRX_MODE_BATCHED;
  cfg0.use_graph = 0;
  const int thr_c[] = {64, 128, 256};
  const int chunk_c[] = {128, 256, 512, 1024, 2048, 4096, 8192, 16384};
  const int reps = 2, ncand = 24;
  const double t_start = now_s();
  // sample size: the whole batch (up to max_walks) unless one run of the
  // default shape takes more than its share of the budget, then halve
  int want = max_walks > 0 ? std::min(max_walks, b->n) : b->n;
  std::unique_ptr<TuneSample> smp;
  for (;;) {
    smp.reset(new TuneSample(b, want));
    double t = 0;
    int rc = time_plan(ctx, &smp->batch, &cfg0, 1, &t);
    if (rc) return rc;
    if (t * (reps + 1) * ncand <= budget_s || want <= 256) break;
    want /= 2;
  }
  int kmax = 1;
  for (const drx_walk& w : smp->walks) kmax = std::max(kmax, w.K);
  const int kp = (int)pow2_ceil((uint32_t)std::min(kmax, kSliceK));
  double best_t = 1e300;
  drx_launch_cfg best_cfg = cfg0;
  for (int thr : thr_c) {
    if (thr < kp) continue;  // a workgroup holds at least one sub-chunk of every prime
    for (int ch : chunk_c) {
      if (best_t < 1e300 && now_s() - t_start > budget_s) break;
      drx_launch_cfg c = cfg0;
      c.threads = thr;
      c.chunk_steps = ch;
      double t = 0;
      const int rc = time_plan(ctx, &smp->batch, &c, reps, &t);
      if (rc == DRX_ERANGE) continue;  // LDS limit for this shape
      if (rc) return rc;
      if (t < best_t) {
        best_t = t;
        best_cfg = c;
      }
    }
  }
  if (best_t >= 1e300) return fail(ctx, DRX_ERANGE, "no launch configuration fits these walks");
  // keep the default shape unless a candidate beats it clearly (run-to-run
  // noise of the concurrent streams is a few percent)
  {
    drx_launch_cfg d = cfg0;
    d.threads = 256;
    d.chunk_steps = 4096;
    double td = 0;
    if (kp <= 256 && time_plan(ctx, &smp->batch, &d, reps, &td) == 0 && td <= best_t * 1.03) {
      best_t = td;
      best_cfg = d;
    }
  }
  // then, at that shape, walks per workgroup (lane groups of at least one
  // prime stripe: fewer sub-chunks per walk, so shorter trees and no table
  // jumps for one-chunk walks) and the prime stripe of wide walks
  auto try_cfg = [&](const drx_launch_cfg& c) -> int {
    double t = 0;
    const int rc = time_plan(ctx, &smp->batch, &c, reps, &t);
    if (rc == DRX_ERANGE) return DRX_OK;
    if (rc) return rc;
    if (t < best_t * 0.97) {  // a clear gain only
      best_t = t;
      best_cfg = c;
    }
    return DRX_OK;
  };
  const int kp_stripe = (int)pow2_ceil((uint32_t)std::min(kmin_walk(smp->walks), kSliceK));
  for (int wpg = 2; wpg * kp_stripe <= best_cfg.threads && now_s() - t_start <= budget_s; wpg *= 2) {
    drx_launch_cfg c = best_cfg;
    c.walks_per_wg = wpg;
    const int rc = try_cfg(c);
    if (rc) return rc;
  }
  for (int sl = 64; kmax > sl && sl < kSliceK && now_s() - t_start <= budget_s; sl *= 2) {
    drx_launch_cfg c = best_cfg;
    c.slice = sl;
    const int rc = try_cfg(c);
    if (rc) return rc;
  }
  best_cfg.use_graph = base ? base->use_graph : 0;
  best_cfg.mode = base ? base->mode : DRX_MODE_BATCHED;
  *best = best_cfg;
  *best_ms = best_t * 1e3;
  return DRX_OK;
}

int drx_plan_stats(const drx_plan* pl, double* s4) {
  if (!pl || !s4) return DRX_EINVAL;
  s4[0] = pl->residue_steps;
  s4[1] = pl->mm_total;
  s4[2] = pl->ma_total;
  s4[3] = pl->ntask;
  return DRX_OK;
}

void drx_plan_destroy(drx_plan* pl) { delete pl; }

int drx_walk_batch(drx_ctx* ctx, const drx_batch* b, const drx_launch_cfg* cfg) {
  drx_plan* pl = nullptr;
  int rc = drx_plan_create(ctx, b, cfg, &pl);
  if (rc) return rc;
  rc = drx_plan_run(pl);
  drx_plan_destroy(pl);
  return rc;
}

int drx_crt(drx_ctx* ctx, const int64_t* residues, int n, int K, int stride, uint32_t* limbs, uint8_t* neg) {
  if (!ctx) return DRX_EINVAL;
  if (n < 0 || K < 1 || stride < K || (n > 0 && (!residues || !limbs || !neg)))
    return fail(ctx, DRX_EINVAL, "bad CRT arguments");
  if (K > (int)ctx->primes.size()) return fail(ctx, DRX_EINVAL, "K=%d exceeds the %zu primes set", K, ctx->primes.size());
  if (n == 0) return DRX_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < K; ++k) {
      const int64_t r = residues[(size_t)i * stride + k];
      if (r < 0 || r >= (int64_t)ctx->primes[k]) return fail(ctx, DRX_EINVAL, "residue row %d prime %d not canonical", i, k);
    }
  const size_t rows = (size_t)n * stride;
  HIPCHK(ctx, ctx->d_tmp0.ensure(rows * 8));
  HIPCHK(ctx, ctx->d_tmp2.ensure(rows * 4));
  HIPCHK(ctx, ctx->d_tmp3.ensure((size_t)n));
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_tmp0.p, residues, rows * 8, hipMemcpyHostToDevice, ctx->stream));
  int rc;
  const int path = crt_path(K, (size_t)n);
  if (path == CRT_MATRIX) {
    const CrtTable* t = nullptr;
    if ((rc = ensure_crt_table(ctx, K, &t))) return rc;
    const int npad = (n + 127) / 128 * 128;
    std::vector<int32_t> ids(n);
    std::iota(ids.begin(), ids.end(), 0);
    HIPCHK(ctx, ctx->d_crt_rows.ensure((size_t)n * 4));
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_crt_rows.p, ids.data(), (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, ctx->d_crt_a.ensure((size_t)4 * npad * t->Kp));
    HIPCHK(ctx, ctx->d_crt_c.ensure((size_t)t->Tp * npad * 4));
    HIPCHK(ctx, ctx->d_crt_k.ensure((size_t)npad * 12));
    CrtArgs c{};
    crt_matrix_args(ctx, t, n, npad, ctx->d_crt_rows.as<int32_t>(), ctx->d_tmp0.as<int64_t>(), (uint64_t)stride,
                    ctx->d_crt_a.as<int8_t>(), ctx->d_crt_c.as<int32_t>(), ctx->d_crt_k.as<int32_t>(),
                    reinterpret_cast<double*>(ctx->d_crt_k.as<char>() + (size_t)npad * 4), ctx->d_tmp2.as<uint32_t>(),
                    ctx->d_tmp3.as<uint8_t>(), c);
    HIPCHK(ctx, launch_crt_matrix(c, ctx->stream));
  } else {
    GarnerArgs g{};
    g.res = ctx->d_tmp0.as<int64_t>();
    g.res_stride = (uint64_t)stride;
    g.n = n;
    g.K = K;
    g.row_k = nullptr;
    g.prime_p = ctx->d_prime_p.as<double>();
    g.prime_pinv = ctx->d_prime_pinv.as<double>();
    g.prime_u = ctx->d_prime_u.as<uint32_t>();
    g.limbs = ctx->d_tmp2.as<uint32_t>();
    g.neg = ctx->d_tmp3.as<uint8_t>();
    if (path == CRT_WAVE) {
      if ((rc = ensure_inv_table(ctx, K))) return rc;
      g.inv_table = ctx->d_inv.as<uint32_t>();
      g.inv_ld = (uint64_t)ctx->kinv;
      HIPCHK(ctx, launch_garner(g, ctx->stream));
    } else {
      if ((rc = ensure_garner_w(ctx, K))) return rc;
      g.w = ctx->d_garner_w.as<double>();
      HIPCHK(ctx, launch_garner_wide(g, ctx->stream));
    }
  }
  HIPCHK(ctx, hipMemcpyAsync(limbs, ctx->d_tmp2.p, rows * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(neg, ctx->d_tmp3.p, (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DRX_OK;
}

int drx_topk(drx_ctx* ctx, const double* values, int n, int k, int32_t* idx) {
  if (!ctx) return DRX_EINVAL;
  if (n < 0 || k < 0 || k > n || (n > 0 && !values) || (k > 0 && !idx))
    return fail(ctx, DRX_EINVAL, "bad top-k arguments");
  if (k == 0) return DRX_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, ctx->d_tmp0.ensure((size_t)n * 8));
  HIPCHK(ctx, ctx->d_tmp1.ensure((size_t)k * 4));
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_tmp0.p, values, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, ctx->d_tmp2.ensure(std::max<size_t>(16, topk_workspace_bytes(n, k))));
  HIPCHK(ctx, launch_topk(ctx->d_tmp0.as<double>(), 1, n, k, ctx->d_tmp1.as<int32_t>(), ctx->d_tmp2.p, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(idx, ctx->d_tmp1.p, (size_t)k * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DRX_OK;
}

int drx_host_alloc(uint64_t bytes, void** out) {
  if (!out) return DRX_EINVAL;
  *out = nullptr;
  if (bytes == 0) bytes = 1;
  return hipHostMalloc(out, bytes, hipHostMallocDefault) == hipSuccess ? DRX_OK : DRX_ENOMEM;
}

void drx_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

int drx_copy_calibrate(drx_ctx* ctx, uint64_t bytes, int reps, double* ms) {
  if (!ctx || !ms || reps <= 0 || bytes < 16 || bytes % 16) return DRX_EINVAL;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  void *src = nullptr, *dst = nullptr;
  hipError_t e = hipMalloc(&src, bytes);
  if (e == hipSuccess) e = hipMalloc(&dst, bytes);
  if (e == hipSuccess) e = hipMemsetAsync(src, 0x5a, bytes, ctx->stream);
  if (e == hipSuccess) e = launch_copy(src, dst, bytes / 16, ctx->stream);  // warm
  if (e == hipSuccess) e = hipEventRecord(ctx->ev[10], ctx->stream);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy(src, dst, bytes / 16, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev[11], ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ev[10], ctx->ev[11]);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  HIPCHK(ctx, e);
  *ms = (double)t / reps;
  return DRX_OK;
}

int drx_valu_rate(drx_ctx* ctx, int op, double* ops_per_s, double* peak_per_s) {
  if (!ctx || !ops_per_s || !peak_per_s || op < 0 || op > 3) return DRX_EINVAL;
  std::lock_guard<std::mutex> lock(ctx->mu);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int ncu = 0, khz = 0;
  HIPCHK(ctx, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
  HIPCHK(ctx, hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, ctx->device));
  HIPCHK(ctx, ctx->d_tmp1.ensure(64));
  const int blocks = 8 * ncu, iters = op == 3 ? 4096 : 16384;
  hipError_t e = launch_rate(op, blocks, 64, ctx->d_tmp1.as<double>(), ctx->stream);  // warm
  if (e == hipSuccess) e = hipEventRecord(ctx->ev[10], ctx->stream);
  if (e == hipSuccess) e = launch_rate(op, blocks, iters, ctx->d_tmp1.as<double>(), ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev[11], ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev[10], ctx->ev[11]);
  HIPCHK(ctx, e);
  const double ops = (double)blocks * 256 * iters * kRateChains;
  *ops_per_s = ops / (ms * 1e-3);
  *peak_per_s = (double)ncu * 64 * khz * 1e3;
  return DRX_OK;
}

int drx_last_timing(const drx_ctx* ctx, double* t6) {
  if (!ctx || !t6) return DRX_EINVAL;
  t6[0] = ctx->timing.walk_ms;
  t6[1] = ctx->timing.combine_ms;
  t6[2] = ctx->timing.shadow_ms;
  t6[3] = ctx->timing.tail_ms;
  t6[4] = ctx->timing.total_ms;
  t6[5] = ctx->timing.walk_launches;
  return DRX_OK;
}

}  // extern "C"
