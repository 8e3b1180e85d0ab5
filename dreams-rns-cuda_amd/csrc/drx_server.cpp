// Continuous-batching walk server.
//
// A background thread runs rounds. Each round (1) admits queued requests into
// free device-resident slots while the slot budget allows (their initial
// matrices and axis tables are built on the device from a small record per
// walk), (2) advances every resident walk with one launch sequence (walk
// kernel per matrix order, ordered combine into the resident state, FP64
// shadow): by one step range of `round_steps` steps, or, while few walks are
// resident, by up to `max_walk_tasks` consecutive step ranges that run side
// by side, and (3) finalises the walks that reached their depth, runs the
// Garner CRT of their p and q, and copies residues, limbs and FP64 results
// into page-locked host memory in the same stream. The context lock is held
// only while a round is built and enqueued, not while the device runs it, so
// submissions never wait for a round. Walk state never leaves the device
// between rounds. A walk's result does not depend on which other walks share
// its rounds or on how its depth is split into rounds and ranges: every round
// multiplies the state by the ordered product of the walk's own M(s), exactly.
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <new>
#include <thread>

#include "drx_ctx.h"
#include "drx_device.h"

using namespace drx;

namespace {

struct SrvReq {
  uint64_t ticket;
  uint64_t epoch;  // context epoch the program id belongs to
  drx_walk w;
  DevAdmit rec;    // initial matrix and combined axis rationals
  double t_submit;
};

struct SrvSlot {
  bool used = false;
  uint64_t ticket = 0;
  uint64_t epoch = 0;
  drx_walk w{};
  int m = 0;
  int64_t step = 0;
  double t_submit = 0, t_admit = 0;
};

struct SrvDone {
  drx_server_result r;
  std::vector<int64_t> res;     // p then q residues
  std::vector<uint32_t> limbs;  // p then q CRT magnitudes
};

template <typename T>
T* host_alloc(size_t n) {
  void* p = nullptr;
  return hipHostMalloc(&p, std::max<size_t>(1, n) * sizeof(T), hipHostMallocDefault) == hipSuccess
             ? static_cast<T*>(p)
             : nullptr;
}

}  // namespace

struct drx_server {
  drx_ctx* ctx = nullptr;
  drx_server_cfg cfg{};
  int kslot = 0;        // residues per entry per slot (the primes set, at most 256)
  int nslots = 0;
  int target_tasks = 0;  // walk-kernel tasks a round aims at when few walks are resident
  std::thread th;
  std::mutex mu;
  std::condition_variable cv;
  bool stop = false;
  int sticky_error = 0;
  std::string sticky_msg;
  std::deque<SrvReq> queue;
  std::vector<SrvSlot> slots;
  std::deque<SrvDone> done;
  uint64_t next_ticket = 1;
  DBuf d_walks, d_tasks, d_spans, d_fspans, d_sspans, d_ranges, d_res, d_fst, d_axr, d_axi, d_axd, d_cp, d_mf, d_fdt,
      d_fout, d_rout, d_fcomp, d_adm, d_rowk, d_limbs, d_neg;
  // page-locked: admission records and CRT row widths in, results out
  DevAdmit* h_adm = nullptr;
  int32_t* h_rowk = nullptr;
  double* h_fcomp = nullptr;
  int64_t* h_rout = nullptr;
  uint32_t* h_limbs = nullptr;
  uint8_t* h_neg = nullptr;
  hipEvent_t ev_done = nullptr;
  // per-round host tables (members so that they outlive the round's copies)
  std::vector<DevWalk> h_walks;
  std::vector<DevTask> ht, hseed;
  std::vector<DevSpan> hs, hf;
  std::vector<DevRange> hr;
  ~drx_server() {
    for (DBuf* d : {&d_walks, &d_tasks, &d_spans, &d_fspans, &d_sspans, &d_ranges, &d_res, &d_fst, &d_axr, &d_axi,
                    &d_axd, &d_cp, &d_mf, &d_fdt, &d_fout, &d_rout, &d_fcomp, &d_adm, &d_rowk, &d_limbs, &d_neg})
      d->release();
    for (void* p : {(void*)h_adm, (void*)h_rowk, (void*)h_fcomp, (void*)h_rout, (void*)h_limbs, (void*)h_neg})
      if (p) (void)hipHostFree(p);
    if (ev_done) (void)hipEventDestroy(ev_done);
  }
};

namespace {

size_t slot_res(const drx_server* s) { return (size_t)16 * s->kslot; }
size_t slot_axr(const drx_server* s) { return (size_t)kMaxDim * 3 * s->kslot; }
// Newton tables: the largest of 2x2 (4 entries x 7), 3x3 (9 entries x 4) and
// 4x4 companions (16 entries x 4) per prime
size_t slot_fdt(const drx_server* s) {
  const size_t a = 4 * (kFdMax2 + 1), b = 16 * (kFdMax3 + 1);
  return (a > b ? a : b) * s->kslot;
}

// The admission record of a walk: the combined rationals of its axes
// (c0 = sn td, c1 = tn sd, den = sd td), checked like the batch path's
// axis tables.
int admit_record(drx_ctx* ctx, const drx_walk& w, int dim, const int64_t* q, DevAdmit* rec) {
  *rec = DevAdmit{};
  rec->init_kind = w.init_kind;
  rec->a0 = w.a0;
  for (int ax = 0; ax < dim; ++ax) {
    const __int128 sn = q[4 * ax], sd = q[4 * ax + 1], tn = q[4 * ax + 2], td = q[4 * ax + 3];
    const __int128 c0 = sn * td, c1 = tn * sd, den = sd * td;
    if (den == 0) return fail(ctx, DRX_EINVAL, "axis %d: zero denominator", ax);
    const __int128 lim = (__int128)1 << 62;
    const __int128 last = c0 + (__int128)w.depth * c1;
    if (c0 > lim || c0 < -lim || c1 > lim || c1 < -lim || den > lim || den < -lim || last > lim || last < -lim)
      return fail(ctx, DRX_ERANGE, "axis %d: shift/trajectory too large for int64", ax);
    rec->c0[ax] = (int64_t)c0;
    rec->c1[ax] = (int64_t)c1;
    rec->den[ax] = (int64_t)den;
  }
  return DRX_OK;
}

// Slot bookkeeping of an admitted walk (its device tables come from k_admit).
void admit_slot(drx_server* s, int slot, const SrvReq& rq) {
  drx_ctx* ctx = s->ctx;
  const HostProgram& hp = ctx->progs[rq.w.prog];
  DevWalk& dw = s->h_walks[slot];
  dw = DevWalk{};
  dw.prog = rq.w.prog;
  dw.K = rq.w.K;
  dw.m = (int)hp.dev.m;
  dw.dim = (int)hp.dev.dim;
  dw.res_off = (uint64_t)slot * slot_res(s);
  dw.fst_off = (uint64_t)slot * kFstMax;
  dw.fdt_off = (uint64_t)slot * slot_fdt(s);
  dw.axr_off = (uint64_t)slot * slot_axr(s);
  dw.axi_off = (uint64_t)slot * kMaxDim * 2;
  dw.axd_off = (uint64_t)slot * kMaxDim;
  dw.target = rq.w.target;
  SrvSlot& sl = s->slots[slot];  // reserved (used) by the server loop
  sl.ticket = rq.ticket;
  sl.w = rq.w;
  sl.m = dw.m;
  sl.step = 0;
  sl.t_submit = rq.t_submit;
  sl.t_admit = now_s();
  sl.epoch = rq.epoch;
}

// Builds and enqueues one round (context lock held): admission of `adm`
// (slots already reserved), the walk pipeline over every resident slot, the
// finalisation and CRT of the walks that finish, and the copies of their
// results into page-locked memory; records ev_done. `finished` gets the
// finishing slots in output-row order.
int round_launch(drx_server* s, const std::vector<SrvReq>& adm, const std::vector<int>& adm_slots,
                 std::vector<int>& finished) {
  drx_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  finished.clear();
  for (size_t j = 0; j < adm.size(); ++j) {
    admit_slot(s, adm_slots[j], adm[j]);
    s->h_adm[j] = adm[j].rec;
    s->h_adm[j].slot = adm_slots[j];
  }
  for (int i = 0; i < s->nslots; ++i)
    if (s->slots[i].used && s->slots[i].epoch != ctx->epoch)
      return fail(ctx, DRX_ESTATE, "programs were released while requests were resident");
  struct TaskRef {
    int slot, m, kind, lclass, shape, prog, K;
    double cost;  // work of one step range: primes x residue-code length
  };
  std::vector<TaskRef> refs;
  for (int i = 0; i < s->nslots; ++i) {
    if (!s->slots[i].used) continue;
    const HostProgram& hp = ctx->progs[s->slots[i].w.prog];
    const int kind = walk_kind(hp, false);
    refs.push_back({i, s->slots[i].m, kind, lds_class(hp, kind), walk_shape(hp, kind), s->slots[i].w.prog,
                    s->slots[i].w.K, (double)s->slots[i].w.K * (double)(1 + hp.pk.code.size())});
  }
  if (refs.empty()) return DRX_OK;
  // within a stage the costliest step ranges first, so that the launch ends
  // on short ones; walks of one program and width stay adjacent (consecutive
  // workgroups then read the same code and constants through the scalar cache)
  std::stable_sort(refs.begin(), refs.end(), [](const TaskRef& a, const TaskRef& b) {
    if (a.m != b.m) return a.m < b.m;
    if (a.kind != b.kind) return a.kind < b.kind;
    if (a.lclass != b.lclass) return a.lclass < b.lclass;
    if (a.shape != b.shape) return a.shape < b.shape;
    if (a.cost != b.cost) return a.cost > b.cost;
    return a.prog < b.prog;
  });
  // step ranges per walk this round: enough walk-kernel tasks to fill the
  // device while few walks are resident, one per walk under load
  const int nres = (int)refs.size();
  const int cap = s->cfg.max_walk_tasks > 0 ? s->cfg.max_walk_tasks : 32;
  const int tpw = std::max(1, std::min(cap, s->target_tasks / nres));
  const int64_t rs = s->cfg.round_steps;
  s->ht.clear();
  s->hseed.clear();
  s->hs.clear();
  s->hf.clear();
  s->hr.clear();
  std::vector<Stage> stages;
  uint64_t mf_total = 0;
  std::vector<int64_t> new_step(s->nslots, 0);
  for (const TaskRef& r : refs) {
    SrvSlot& sl = s->slots[r.slot];
    const DevWalk& dw = s->h_walks[r.slot];
    if (stages.empty() || stages.back().m != r.m || stages.back().kind != r.kind || stages.back().lclass != r.lclass) {
      Stage g;
      g.m = r.m;
      g.kind = r.kind;
      g.lclass = r.lclass;
      g.task_begin = g.task_end = (int)s->ht.size();
      g.span_begin = g.span_end = (int)s->hs.size();
      g.seed_begin = g.seed_end = (int)s->hseed.size();
      g.range_begin = g.range_end = (int)s->hr.size();
      stages.push_back(g);
    }
    Stage& g = stages.back();
    const int threads = std::max<int>(s->cfg.threads > 0 ? s->cfg.threads : 256, (int)pow2_ceil(dw.K));
    g.threads = std::max(g.threads, threads);
    g.seed_kp = std::max<int>(g.seed_kp, (int)pow2_ceil(dw.K));
    DevSpan sp{};
    sp.walk = r.slot;
    sp.task_begin = (int)s->ht.size();
    sp.out_row = r.slot;
    mf_total = (mf_total + 1) & ~(uint64_t)1;  // 16-byte aligned M_f; a walk's ranges stay contiguous
    int64_t s0 = sl.step;
    for (int k = 0; k < tpw && (k == 0 || s0 < sl.w.depth); ++k) {
      DevTask t{};
      t.walk = r.slot;
      t.shape = r.shape;
      t.s_begin = s0;
      t.s_end = std::min<int64_t>(sl.w.depth, s0 + rs);
      t.cp_off = (uint64_t)s->ht.size() * slot_res(s);
      t.k0 = 0;
      t.kn = dw.K;
      t.mf_off = mf_total;
      mf_total += (uint64_t)(t.s_end - t.s_begin) * r.m * r.m;
      for (int64_t q = t.s_begin; q < t.s_end; q += kMfRange) {
        DevRange rg{};
        rg.task = (int)s->ht.size();
        rg.s_begin = q;
        rg.s_end = std::min<int64_t>(t.s_end, q + kMfRange);
        s->hr.push_back(rg);
      }
      if (r.kind != WALK_INTERP && t.s_begin == 0) s->hseed.push_back(t);  // Newton tables once, at step 0
      s->ht.push_back(t);
      s0 = t.s_end;
    }
    new_step[r.slot] = s0;
    sp.task_end = (int)s->ht.size();
    s->hs.push_back(sp);
    g.task_end = (int)s->ht.size();
    g.span_end = (int)s->hs.size();
    g.range_end = (int)s->hr.size();
    g.seed_end = (int)s->hseed.size();
    if (s0 >= sl.w.depth) {
      sp.out_row = (int)s->hf.size();
      s->hf.push_back(sp);
      finished.push_back(r.slot);
    }
  }
  const size_t lds_limit = device_lds_limit(ctx);
  for (Stage& g : stages) {
    // halve the walk workgroup (never below the widest prime stripe) until
    // its LDS fits; the sub-chunk split changes, the product does not
    for (;;) {
      size_t kp_max = 1;
      g.lds = 0;
      for (int ti = g.task_begin; ti < g.task_end; ++ti) {
        const DevWalk& dw = s->h_walks[s->ht[ti].walk];
        g.lds = std::max(g.lds, walk_lds_bytes(ctx->progs[dw.prog].dev, g.kind, g.m, (uint32_t)dw.K, g.threads));
        kp_max = std::max<size_t>(kp_max, pow2_ceil((uint32_t)dw.K));
      }
      if (g.lds <= lds_limit || (size_t)g.threads / 2 < kp_max) break;
      g.threads /= 2;
    }
    for (int ti = g.task_begin; ti < g.task_end; ++ti) {
      const DevWalk& dw = s->h_walks[s->ht[ti].walk];
      const DevProgram& dp = ctx->progs[dw.prog].dev;
      g.mf_lds = std::max(g.mf_lds, mf_lds_bytes(dp));
      if (g.kind != WALK_INTERP)
        g.seed_lds = std::max(g.seed_lds, seed_lds_bytes(dp, g.seed_kp, fd_seed_threads(g.m, g.seed_kp)));
    }
    if (std::max(g.lds, std::max(g.mf_lds, g.seed_lds)) > lds_limit)
      return fail(ctx, DRX_ERANGE, "server round needs more LDS than a workgroup may have");
  }
  merge_lds_classes(stages);
  const size_t nt = s->ht.size(), nh = s->hs.size(), nf = s->hf.size();
  DRX_HIPCHK(ctx, s->d_tasks.ensure(nt * sizeof(DevTask)));
  DRX_HIPCHK(ctx, s->d_spans.ensure(nh * sizeof(DevSpan)));
  DRX_HIPCHK(ctx, s->d_fspans.ensure(std::max<size_t>(1, nf) * sizeof(DevSpan)));
  DRX_HIPCHK(ctx, s->d_sspans.ensure(std::max<size_t>(1, s->hseed.size()) * sizeof(DevTask)));
  DRX_HIPCHK(ctx, s->d_ranges.ensure(std::max<size_t>(1, s->hr.size()) * sizeof(DevRange)));
  DRX_HIPCHK(ctx, s->d_cp.ensure(nt * slot_res(s) * 8));
  DRX_HIPCHK(ctx, s->d_mf.ensure((std::max<uint64_t>(16, mf_total) + kMfSlack) * 8));
  DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_walks.p, s->h_walks.data(), s->h_walks.size() * sizeof(DevWalk), hipMemcpyHostToDevice, st));
  DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_tasks.p, s->ht.data(), nt * sizeof(DevTask), hipMemcpyHostToDevice, st));
  DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_spans.p, s->hs.data(), nh * sizeof(DevSpan), hipMemcpyHostToDevice, st));
  if (nf) DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_fspans.p, s->hf.data(), nf * sizeof(DevSpan), hipMemcpyHostToDevice, st));
  if (!s->hseed.empty())
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_sspans.p, s->hseed.data(), s->hseed.size() * sizeof(DevTask), hipMemcpyHostToDevice, st));
  if (!s->hr.empty())
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_ranges.p, s->hr.data(), s->hr.size() * sizeof(DevRange), hipMemcpyHostToDevice, st));
  if (!adm.empty())
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_adm.p, s->h_adm, adm.size() * sizeof(DevAdmit), hipMemcpyHostToDevice, st));
  int rc = upload_programs(ctx);
  if (rc) return rc;
  KArgs a{};
  a.tasks = s->d_tasks.as<DevTask>();
  a.walks = s->d_walks.as<DevWalk>();
  a.progs = ctx->d_progs.as<DevProgram>();
  a.spans = s->d_spans.as<DevSpan>();
  a.ranges = s->d_ranges.as<DevRange>();
  a.code = ctx->d_code.as<uint32_t>();
  a.cres = ctx->d_cres.as<double>();
  a.cflt = ctx->d_cflt.as<double>();
  a.prime_p = ctx->d_prime_p.as<double>();
  a.prime_pinv = ctx->d_prime_pinv.as<double>();
  a.prime_e = ctx->d_prime_e.as<uint32_t>();
  a.invs = ctx->d_invs.as<double>();
  a.axr = s->d_axr.as<double>();
  a.axi = s->d_axi.as<int64_t>();
  a.axd = s->d_axd.as<double>();
  a.axr_w = s->d_axr.as<double>();
  a.axi_w = s->d_axi.as<int64_t>();
  a.axd_w = s->d_axd.as<double>();
  a.res_in = s->d_res.as<double>();
  a.fst_in = s->d_fst.as<double>();
  a.res = s->d_res.as<double>();
  a.fst = s->d_fst.as<double>();
  a.cp = s->d_cp.as<double>();
  a.mf = s->d_mf.as<double>();
  a.fdt = s->d_fdt.as<double>();
  a.fout = s->d_fout.as<double>();
  a.rout = s->d_rout.as<int64_t>();
  a.pout = nullptr;
  a.fcompact = s->d_fcomp.as<double>();
  a.ffull = nullptr;
  a.out_ld = (uint64_t)s->kslot;
  if (!adm.empty()) DRX_HIPCHK(ctx, launch_admit(a, s->d_adm.as<DevAdmit>(), (int)adm.size(), st));
  for (Stage& g : stages) g.win_range = {g.range_begin, g.range_end};  // one window per round
  rc = launch_stages(ctx, a, stages, s->d_sspans.as<DevTask>(), 1, false, INT64_MAX, 1);
  if (!rc) rc = join_stages(ctx, false);
  if (rc) return rc;
  if (nf) {
    KArgs fa = a;
    fa.spans = s->d_fspans.as<DevSpan>();
    DRX_HIPCHK(ctx, launch_finalize(fa, (int)nf, 0, st));
    // CRT of the finished p and q rows (at most 256 primes: the wavefront Garner kernel)
    int kmax = 1;
    for (size_t j = 0; j < nf; ++j) {
      const int K = s->slots[finished[j]].w.K;
      s->h_rowk[2 * j] = s->h_rowk[2 * j + 1] = K;
      kmax = std::max(kmax, K);
    }
    if ((rc = ensure_inv_table(ctx, kmax))) return rc;
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_rowk.p, s->h_rowk, nf * 2 * 4, hipMemcpyHostToDevice, st));
    GarnerArgs g{};
    g.res = s->d_rout.as<int64_t>();
    g.res_stride = (uint64_t)s->kslot;
    g.n = (int)(2 * nf);
    g.K = kmax;
    g.row_k = s->d_rowk.as<int32_t>();
    g.prime_p = ctx->d_prime_p.as<double>();
    g.prime_pinv = ctx->d_prime_pinv.as<double>();
    g.prime_u = ctx->d_prime_u.as<uint32_t>();
    g.inv_table = ctx->d_inv.as<uint32_t>();
    g.inv_ld = (uint64_t)ctx->kinv;
    g.limbs = s->d_limbs.as<uint32_t>();
    g.neg = s->d_neg.as<uint8_t>();
    DRX_HIPCHK(ctx, launch_garner(g, st));
    const size_t row = (size_t)2 * s->kslot;
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->h_fcomp, s->d_fcomp.p, nf * 6 * 8, hipMemcpyDeviceToHost, st));
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->h_rout, s->d_rout.p, nf * row * 8, hipMemcpyDeviceToHost, st));
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->h_limbs, s->d_limbs.p, nf * row * 4, hipMemcpyDeviceToHost, st));
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->h_neg, s->d_neg.p, nf * 2, hipMemcpyDeviceToHost, st));
  }
  DRX_HIPCHK(ctx, hipEventRecord(s->ev_done, st));
  for (const TaskRef& r : refs) s->slots[r.slot].step = new_step[r.slot];
  return DRX_OK;
}

// After ev_done: results of the finished walks into the done queue.
void harvest(drx_server* s, const std::vector<int>& finished) {
  const double t_done = now_s();
  const size_t row = (size_t)2 * s->kslot;
  std::lock_guard<std::mutex> lk(s->mu);
  for (size_t j = 0; j < finished.size(); ++j) {
    SrvSlot& sl = s->slots[finished[j]];
    SrvDone d;
    d.r = drx_server_result{};
    d.r.ticket = sl.ticket;
    d.r.K = sl.w.K;
    d.r.status = 0;
    d.r.t_submit = sl.t_submit;
    d.r.t_admit = sl.t_admit;
    d.r.t_done = t_done;
    const double* f6 = s->h_fcomp + 6 * j;
    d.r.out.p_float = f6[0];
    d.r.out.q_float = f6[1];
    d.r.out.log_scale = f6[2];
    d.r.out.delta_float = f6[3];
    d.r.out.log_q = f6[4];
    d.r.out.reserved = 0;
    d.r.neg[0] = s->h_neg[2 * j];
    d.r.neg[1] = s->h_neg[2 * j + 1];
    const int K = sl.w.K;
    const int64_t* src = s->h_rout + j * row;
    d.res.assign(src, src + K);
    d.res.insert(d.res.end(), src + s->kslot, src + s->kslot + K);
    const uint32_t* ls = s->h_limbs + j * row;
    d.limbs.assign(ls, ls + K);
    d.limbs.insert(d.limbs.end(), ls + s->kslot, ls + s->kslot + K);
    s->done.push_back(std::move(d));
    sl.used = false;
  }
}

void fail_resident(drx_server* s, int rc) {
  std::lock_guard<std::mutex> lk(s->mu);
  s->sticky_error = rc;
  s->sticky_msg = s->ctx->err;
  for (int i = 0; i < s->nslots; ++i) {
    SrvSlot& sl = s->slots[i];
    if (!sl.used) continue;
    SrvDone d{};
    d.r.ticket = sl.ticket;
    d.r.status = rc;
    d.r.t_submit = sl.t_submit;
    d.r.t_admit = sl.t_admit;
    d.r.t_done = now_s();
    s->done.push_back(std::move(d));
    sl.used = false;
  }
}

void server_loop(drx_server* s) {
  std::vector<int> finished;
  for (;;) {
    std::vector<SrvReq> admitted;
    std::vector<int> free_slots;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] {
        if (s->stop) return true;
        if (!s->queue.empty()) return true;
        for (const SrvSlot& sl : s->slots)
          if (sl.used) return true;
        return false;
      });
      if (s->stop) return;
      for (int i = 0; i < s->nslots && !s->queue.empty(); ++i) {
        if (s->slots[i].used) continue;
        admitted.push_back(std::move(s->queue.front()));
        s->queue.pop_front();
        free_slots.push_back(i);
        s->slots[i].used = true;  // reserved
      }
    }
    int rc;
    {
      std::lock_guard<std::mutex> ctx_lock(s->ctx->mu);
      (void)hipSetDevice(s->ctx->device);
      // requests whose programs were released since submission fail alone
      std::vector<SrvReq> ok;
      std::vector<int> ok_slots;
      for (size_t j = 0; j < admitted.size(); ++j) {
        if (admitted[j].epoch == s->ctx->epoch) {
          ok.push_back(std::move(admitted[j]));
          ok_slots.push_back(free_slots[j]);
          continue;
        }
        SrvDone d{};
        d.r.ticket = admitted[j].ticket;
        d.r.status = DRX_ESTATE;
        d.r.t_submit = admitted[j].t_submit;
        d.r.t_done = now_s();
        std::lock_guard<std::mutex> lk(s->mu);
        s->done.push_back(std::move(d));
        s->slots[free_slots[j]].used = false;
      }
      rc = round_launch(s, ok, ok_slots, finished);
    }
    if (!rc) {
      const hipError_t e = hipEventSynchronize(s->ev_done);
      if (e != hipSuccess) rc = fail(s->ctx, DRX_EHIP, "server round: %s", hipGetErrorString(e));
    }
    if (rc) {
      fail_resident(s, rc);  // fail every resident request; keep serving new ones
      continue;
    }
    harvest(s, finished);
  }
}

}  // namespace

extern "C" {

int drx_server_create(drx_ctx* ctx, const drx_server_cfg* cfg, drx_server** out) {
  if (!ctx) return DRX_EINVAL;
  if (!cfg || !out) return fail(ctx, DRX_EINVAL, "bad server arguments");
  *out = nullptr;
  if (ctx->primes.empty()) return fail(ctx, DRX_ESTATE, "set primes before creating a server");
  if (cfg->round_steps < 1) return fail(ctx, DRX_EINVAL, "round_steps must be >= 1");
  if (cfg->threads != 0 && (cfg->threads < 32 || cfg->threads > 256 || (cfg->threads & (cfg->threads - 1))))
    return fail(ctx, DRX_EINVAL, "threads must be 0 or a power of two in [32, 256]");
  if (cfg->max_walk_tasks < 0 || cfg->max_walk_tasks > 64)
    return fail(ctx, DRX_EINVAL, "max_walk_tasks must be in [0, 64]");
  drx_server* s = new (std::nothrow) drx_server();
  if (!s) return fail(ctx, DRX_ENOMEM, "out of host memory");
  s->ctx = ctx;
  s->cfg = *cfg;
  s->kslot = std::min<int>((int)ctx->primes.size(), kSliceK);  // a served walk takes at most 256 primes
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || ncu < 1)
    ncu = 64;
  const int64_t budget = cfg->mem_budget > 0 ? cfg->mem_budget : ((int64_t)1 << 30);
  // scratch of one extra step range (M_f, chunk product, task and ranges)
  const int64_t per_task = (int64_t)8 * 16 * cfg->round_steps + (int64_t)8 * slot_res(s) + (int64_t)sizeof(DevTask) +
                           (int64_t)sizeof(DevRange) * ((cfg->round_steps + kMfRange - 1) / kMfRange);
  // resident state + per-round scratch of one slot
  const int64_t per_slot =
      (int64_t)8 * (slot_res(s) * 2 + kFstMax + slot_axr(s) + slot_fdt(s) + kMaxDim * 3 + 6 + 2 * s->kslot) +
      (int64_t)4 * 2 * s->kslot + 2 + 8 + (int64_t)sizeof(DevAdmit) + (int64_t)8 * 16 * cfg->round_steps +
      (int64_t)sizeof(DevWalk) + 2 * sizeof(DevTask) + 3 * sizeof(DevSpan) +
      (int64_t)sizeof(DevRange) * ((cfg->round_steps + kMfRange - 1) / kMfRange);
  int64_t cap = budget / per_slot;
  if (cfg->max_slots > 0) cap = std::min<int64_t>(cap, cfg->max_slots);
  if (cap < 1) {
    delete s;
    return fail(ctx, DRX_ENOMEM, "memory budget too small for one slot");
  }
  s->nslots = (int)cap;
  // the extra step ranges of lightly loaded rounds use what the slots leave
  // of the budget, up to 16 walk-kernel tasks per CU: a launch of only a
  // few times the resident workgroups loses much of its time to its tail
  s->target_tasks = (int)std::min<int64_t>(16 * (int64_t)ncu, cap + (budget - cap * per_slot) / per_task);
  s->slots.resize(s->nslots);
  s->h_walks.resize(s->nslots);
  const size_t ns = (size_t)s->nslots;
  if (s->d_res.ensure(ns * slot_res(s) * 8) || s->d_fst.ensure(ns * kFstMax * 8) || s->d_axr.ensure(ns * slot_axr(s) * 8) ||
      s->d_fdt.ensure(ns * slot_fdt(s) * 8) || s->d_axi.ensure(ns * kMaxDim * 2 * 8) || s->d_axd.ensure(ns * kMaxDim * 8) ||
      s->d_fout.ensure(ns * 6 * 8) || s->d_rout.ensure(ns * 2 * s->kslot * 8) || s->d_fcomp.ensure(ns * 6 * 8) ||
      s->d_walks.ensure(ns * sizeof(DevWalk)) || s->d_adm.ensure(ns * sizeof(DevAdmit)) ||
      s->d_rowk.ensure(ns * 2 * 4) || s->d_limbs.ensure(ns * 2 * s->kslot * 4) || s->d_neg.ensure(ns * 2)) {
    delete s;
    return fail(ctx, DRX_ENOMEM, "device allocation for %d slots failed", (int)ns);
  }
  s->h_adm = host_alloc<DevAdmit>(ns);
  s->h_rowk = host_alloc<int32_t>(2 * ns);
  s->h_fcomp = host_alloc<double>(6 * ns);
  s->h_rout = host_alloc<int64_t>(2 * ns * s->kslot);
  s->h_limbs = host_alloc<uint32_t>(2 * ns * s->kslot);
  s->h_neg = host_alloc<uint8_t>(2 * ns);
  if (!s->h_adm || !s->h_rowk || !s->h_fcomp || !s->h_rout || !s->h_limbs || !s->h_neg ||
      hipEventCreateWithFlags(&s->ev_done, hipEventDisableTiming) != hipSuccess) {
    delete s;
    return fail(ctx, DRX_ENOMEM, "page-locked buffers or event for %d slots failed", (int)ns);
  }
  s->th = std::thread(server_loop, s);
  *out = s;
  return DRX_OK;
}

int drx_server_submit(drx_server* s, const drx_walk* w, const int64_t* axis_vals, uint64_t* ticket) {
  if (!s || !w || !ticket) return DRX_EINVAL;
  const double t_submit = now_s();
  drx_ctx* ctx = s->ctx;
  SrvReq rq;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (w->prog < 0 || w->prog >= (int)ctx->progs.size()) return fail(ctx, DRX_EINVAL, "bad program id %d", w->prog);
    const DevProgram& dp = ctx->progs[w->prog].dev;
    if (w->mode != DRX_WALK_MATRIX || ctx->progs[w->prog].mode != DRX_WALK_MATRIX)
      return fail(ctx, DRX_EINVAL, "the server runs matrix walks");
    if (w->K < 1 || w->K > s->kslot || w->K > (int)dp.kcap || pow2_ceil(w->K) > 256)
      return fail(ctx, DRX_EINVAL, "bad K=%d", w->K);
    if (w->K > ctx->kwalk) return fail(ctx, DRX_ERANGE, "K=%d: walks need primes in (2^30, 2^31)", w->K);
    if (w->depth < 0 || w->depth > ((int64_t)1 << 30)) return fail(ctx, DRX_EINVAL, "bad depth");
    if (w->init_kind != DRX_INIT_PCF && w->init_kind != DRX_INIT_IDENTITY)
      return fail(ctx, DRX_EINVAL, "server walks take PCF or identity initial matrices");
    if (w->init_kind == DRX_INIT_PCF && dp.m != 2) return fail(ctx, DRX_EINVAL, "PCF initial matrix needs m=2");
    if (dp.dim > 0 && !axis_vals) return fail(ctx, DRX_EINVAL, "axis values missing");
    const int rc = admit_record(ctx, *w, (int)dp.dim, axis_vals, &rq.rec);
    if (rc) return rc;
    rq.epoch = ctx->epoch;
  }
  rq.w = *w;
  rq.t_submit = t_submit;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    rq.ticket = s->next_ticket++;
    *ticket = rq.ticket;
    s->queue.push_back(std::move(rq));
  }
  s->cv.notify_one();
  return DRX_OK;
}

int drx_server_poll(drx_server* s, drx_server_result* res, int64_t* residues, uint32_t* limbs, int stride, int max,
                    int* n_done) {
  if (!s || !n_done || max < 0 || (max > 0 && !res)) return DRX_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  int n = 0;
  while (n < max && !s->done.empty()) {
    SrvDone& d = s->done.front();
    if ((residues || limbs) && !d.res.empty() && stride < d.r.K)
      return fail(s->ctx, DRX_EINVAL, "residue stride %d < K=%d", stride, d.r.K);
    if (residues && !d.res.empty()) {
      int64_t* dst = residues + (size_t)n * 2 * stride;
      std::copy(d.res.begin(), d.res.begin() + d.r.K, dst);
      std::copy(d.res.begin() + d.r.K, d.res.end(), dst + stride);
    }
    if (limbs && !d.limbs.empty()) {
      uint32_t* dst = limbs + (size_t)n * 2 * stride;
      std::fill(dst, dst + (size_t)2 * stride, 0u);
      std::copy(d.limbs.begin(), d.limbs.begin() + d.r.K, dst);
      std::copy(d.limbs.begin() + d.r.K, d.limbs.end(), dst + stride);
    }
    res[n++] = d.r;
    s->done.pop_front();
  }
  *n_done = n;
  return DRX_OK;
}

int drx_server_pending(drx_server* s, int* n_queued, int* n_resident) {
  if (!s) return DRX_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  if (n_queued) *n_queued = (int)s->queue.size();
  if (n_resident) {
    int r = 0;
    for (const SrvSlot& sl : s->slots) r += sl.used;
    *n_resident = r;
  }
  return DRX_OK;
}

int drx_server_capacity(drx_server* s, int* n_slots) {
  if (!s || !n_slots) return DRX_EINVAL;
  *n_slots = s->nslots;
  return DRX_OK;
}

void drx_server_destroy(drx_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->th.joinable()) s->th.join();
  if (s->ctx && s->ev_done) (void)hipEventSynchronize(s->ev_done);
  delete s;
}

}  // extern "C"
