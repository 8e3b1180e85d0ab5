// Continuous-batching walk server.
//
// A background thread runs rounds. Each round (1) admits queued requests into
// free device-resident slots while the slot budget allows, (2) advances every
// resident walk by up to `round_steps` steps with one launch sequence (walk
// kernel per matrix order, ordered combine into the resident state, FP64
// shadow), and (3) finalises the walks that reached their depth and returns
// them to the host. Walk state (P residues, P_f, log_scale) never leaves the
// device between rounds. A walk's result does not depend on which other walks
// share its rounds or on the round length: every round multiplies the state by
// the ordered product of the walk's own M(s), exactly.
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
  std::vector<int64_t> axis;
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
  std::vector<int64_t> res;  // p then q residues
};

}  // namespace

struct drx_server {
  drx_ctx* ctx = nullptr;
  drx_server_cfg cfg{};
  int kslot = 0;        // residues per entry per slot (the primes set, at most 256)
  int nslots = 0;
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
      d_fout, d_rout, d_fcomp;
  std::vector<DevWalk> h_walks;
  ~drx_server() {
    for (DBuf* d : {&d_walks, &d_tasks, &d_spans, &d_fspans, &d_sspans, &d_ranges, &d_res, &d_fst, &d_axr, &d_axi,
                    &d_axd, &d_cp, &d_mf, &d_fdt, &d_fout, &d_rout, &d_fcomp})
      d->release();
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

// Host tables for a newly admitted walk, copied into its slot.
int admit(drx_server* s, int slot, SrvReq& rq) {
  drx_ctx* ctx = s->ctx;
  const drx_walk& w = rq.w;
  if (rq.epoch != ctx->epoch) return fail(ctx, DRX_ESTATE, "programs were released after this request was submitted");
  const HostProgram& hp = ctx->progs[w.prog];
  const int m = (int)hp.dev.m, dim = (int)hp.dev.dim, K = w.K;
  std::vector<double> res((size_t)m * m * K), f(fst_size(m)), axr((size_t)dim * 3 * K), axd(dim);
  std::vector<int64_t> axi((size_t)dim * 2);
  int rc = init_tables(ctx, w, m, nullptr, nullptr, res.data(), f.data());
  if (rc) return rc;
  if (dim > 0 && (rc = axis_tables(ctx, rq.axis.data(), dim, K, w.depth, axr.data(), axi.data(), axd.data())))
    return rc;
  DevWalk& dw = s->h_walks[slot];
  dw.prog = w.prog;
  dw.K = K;
  dw.m = m;
  dw.dim = dim;
  dw.res_off = (uint64_t)slot * slot_res(s);
  dw.fst_off = (uint64_t)slot * kFstMax;
  dw.fdt_off = (uint64_t)slot * slot_fdt(s);
  dw.axr_off = (uint64_t)slot * slot_axr(s);
  dw.axi_off = (uint64_t)slot * kMaxDim * 2;
  dw.axd_off = (uint64_t)slot * kMaxDim;
  dw.target = w.target;
  hipStream_t st = ctx->stream;
  DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_res.as<double>() + dw.res_off, res.data(), res.size() * 8, hipMemcpyHostToDevice, st));
  DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_fst.as<double>() + dw.fst_off, f.data(), f.size() * 8, hipMemcpyHostToDevice, st));
  if (dim > 0) {
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_axr.as<double>() + dw.axr_off, axr.data(), axr.size() * 8, hipMemcpyHostToDevice, st));
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_axi.as<int64_t>() + dw.axi_off, axi.data(), axi.size() * 8, hipMemcpyHostToDevice, st));
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_axd.as<double>() + dw.axd_off, axd.data(), axd.size() * 8, hipMemcpyHostToDevice, st));
  }
  DRX_HIPCHK(ctx, hipStreamSynchronize(st));  // pageable sources
  SrvSlot& sl = s->slots[slot];
  sl.used = true;
  sl.ticket = rq.ticket;
  sl.w = w;
  sl.m = m;
  sl.step = 0;
  sl.t_submit = rq.t_submit;
  sl.t_admit = now_s();
  sl.epoch = rq.epoch;
  return DRX_OK;
}

// One round over the resident slots. Returns the slots finished this round.
int round_once(drx_server* s, std::vector<int>& finished) {
  drx_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  finished.clear();
  for (int i = 0; i < s->nslots; ++i)
    if (s->slots[i].used && s->slots[i].epoch != ctx->epoch)
      return fail(ctx, DRX_ESTATE, "programs were released while requests were resident");
  struct TaskRef {
    int slot, m, kind, lclass, shape;
  };
  std::vector<TaskRef> refs;
  for (int i = 0; i < s->nslots; ++i) {
    if (!s->slots[i].used) continue;
    const HostProgram& hp = ctx->progs[s->slots[i].w.prog];
    const int kind = walk_kind(hp, false);
    refs.push_back({i, s->slots[i].m, kind, lds_class(hp, kind), walk_shape(hp, kind)});
  }
  if (refs.empty()) return DRX_OK;
  std::stable_sort(refs.begin(), refs.end(), [](const TaskRef& a, const TaskRef& b) {
    if (a.m != b.m) return a.m < b.m;
    if (a.kind != b.kind) return a.kind < b.kind;
    if (a.lclass != b.lclass) return a.lclass < b.lclass;
    return a.shape < b.shape;
  });
  std::vector<DevTask> ht, hseed;
  std::vector<DevSpan> hs, hf;
  std::vector<DevRange> hr;
  std::vector<Stage> stages;
  uint64_t mf_total = 0;
  for (const TaskRef& r : refs) {
    SrvSlot& sl = s->slots[r.slot];
    const DevWalk& dw = s->h_walks[r.slot];
    const DevProgram& dp = ctx->progs[dw.prog].dev;
    DevTask t{};
    t.walk = r.slot;
    t.shape = r.shape;
    t.s_begin = sl.step;
    t.s_end = std::min<int64_t>(sl.w.depth, sl.step + s->cfg.round_steps);
    t.cp_off = (uint64_t)ht.size() * slot_res(s);
    t.k0 = 0;
    t.kn = dw.K;
    mf_total = (mf_total + 1) & ~(uint64_t)1;  // 16-byte aligned M_f
    t.mf_off = mf_total;
    mf_total += (uint64_t)(t.s_end - t.s_begin) * r.m * r.m;
    if (stages.empty() || stages.back().m != r.m || stages.back().kind != r.kind || stages.back().lclass != r.lclass) {
      Stage g;
      g.m = r.m;
      g.kind = r.kind;
      g.lclass = r.lclass;
      g.task_begin = g.task_end = g.span_begin = g.span_end = (int)ht.size();
      g.seed_begin = g.seed_end = (int)hseed.size();
      g.range_begin = g.range_end = (int)hr.size();
      stages.push_back(g);
    }
    Stage& g = stages.back();
    const int threads = std::max<int>(s->cfg.threads > 0 ? s->cfg.threads : 256, (int)pow2_ceil(dw.K));
    g.threads = std::max(g.threads, threads);
    g.seed_kp = std::max<int>(g.seed_kp, (int)pow2_ceil(dw.K));
    DevSpan sp{};
    sp.walk = r.slot;
    sp.task_begin = (int)ht.size();
    sp.task_end = (int)ht.size() + 1;
    sp.out_row = r.slot;
    for (int64_t q = t.s_begin; q < t.s_end; q += kMfRange) {
      DevRange rg{};
      rg.task = (int)ht.size();
      rg.s_begin = q;
      rg.s_end = std::min<int64_t>(t.s_end, q + kMfRange);
      hr.push_back(rg);
    }
    ht.push_back(t);
    g.task_end = g.span_end = (int)ht.size();
    g.range_end = (int)hr.size();
    hs.push_back(sp);
    if (r.kind != WALK_INTERP && sl.step == 0) {  // Newton tables once, on the walk's first round
      hseed.push_back(t);
      g.seed_end = (int)hseed.size();
    }
    if (t.s_end >= sl.w.depth) {
      sp.out_row = (int)hf.size();
      hf.push_back(sp);
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
        const DevWalk& dw = s->h_walks[ht[ti].walk];
        g.lds = std::max(g.lds, walk_lds_bytes(ctx->progs[dw.prog].dev, g.kind, g.m, (uint32_t)dw.K, g.threads));
        kp_max = std::max<size_t>(kp_max, pow2_ceil((uint32_t)dw.K));
      }
      if (g.lds <= lds_limit || (size_t)g.threads / 2 < kp_max) break;
      g.threads /= 2;
    }
    for (int ti = g.task_begin; ti < g.task_end; ++ti) {
      const DevWalk& dw = s->h_walks[ht[ti].walk];
      const DevProgram& dp = ctx->progs[dw.prog].dev;
      g.mf_lds = std::max(g.mf_lds, mf_lds_bytes(dp));
      if (g.kind != WALK_INTERP)
        g.seed_lds = std::max(g.seed_lds, seed_lds_bytes(dp, g.seed_kp, fd_seed_threads(g.m, g.seed_kp)));
    }
    if (std::max(g.lds, std::max(g.mf_lds, g.seed_lds)) > lds_limit)
      return fail(ctx, DRX_ERANGE, "server round needs more LDS than a workgroup may have");
  }
  merge_lds_classes(stages);
  DRX_HIPCHK(ctx, s->d_tasks.ensure(ht.size() * sizeof(DevTask)));
  DRX_HIPCHK(ctx, s->d_spans.ensure(hs.size() * sizeof(DevSpan)));
  DRX_HIPCHK(ctx, s->d_fspans.ensure(std::max<size_t>(1, hf.size()) * sizeof(DevSpan)));
  DRX_HIPCHK(ctx, s->d_sspans.ensure(std::max<size_t>(1, hseed.size()) * sizeof(DevTask)));
  DRX_HIPCHK(ctx, s->d_ranges.ensure(std::max<size_t>(1, hr.size()) * sizeof(DevRange)));
  DRX_HIPCHK(ctx, s->d_cp.ensure(ht.size() * slot_res(s) * 8));
  DRX_HIPCHK(ctx, s->d_mf.ensure(std::max<uint64_t>(16, mf_total) * 8));
  DRX_HIPCHK(ctx, s->d_walks.ensure(s->h_walks.size() * sizeof(DevWalk)));
  DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_walks.p, s->h_walks.data(), s->h_walks.size() * sizeof(DevWalk), hipMemcpyHostToDevice, st));
  DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_tasks.p, ht.data(), ht.size() * sizeof(DevTask), hipMemcpyHostToDevice, st));
  DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_spans.p, hs.data(), hs.size() * sizeof(DevSpan), hipMemcpyHostToDevice, st));
  if (!hf.empty())
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_fspans.p, hf.data(), hf.size() * sizeof(DevSpan), hipMemcpyHostToDevice, st));
  if (!hseed.empty())
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_sspans.p, hseed.data(), hseed.size() * sizeof(DevTask), hipMemcpyHostToDevice, st));
  if (!hr.empty())
    DRX_HIPCHK(ctx, hipMemcpyAsync(s->d_ranges.p, hr.data(), hr.size() * sizeof(DevRange), hipMemcpyHostToDevice, st));
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
  for (Stage& g : stages) g.win_range = {g.range_begin, g.range_end};  // one window per round
  rc = launch_stages(ctx, a, stages, s->d_sspans.as<DevTask>(), 1, false, INT64_MAX, 1);
  if (!rc) rc = join_stages(ctx, false);
  if (rc) return rc;
  if (!hf.empty()) {
    KArgs fa = a;
    fa.spans = s->d_fspans.as<DevSpan>();
    DRX_HIPCHK(ctx, launch_finalize(fa, (int)hf.size(), 0, st));
  }
  DRX_HIPCHK(ctx, hipStreamSynchronize(st));
  for (const TaskRef& r : refs) {
    SrvSlot& sl = s->slots[r.slot];
    sl.step = std::min<int64_t>(sl.w.depth, sl.step + s->cfg.round_steps);
  }
  return DRX_OK;
}

int harvest(drx_server* s, const std::vector<int>& finished) {
  drx_ctx* ctx = s->ctx;
  const size_t nf = finished.size();
  if (nf == 0) return DRX_OK;
  std::vector<double> f6(nf * 6);
  std::vector<int64_t> rr(nf * 2 * (size_t)s->kslot);
  DRX_HIPCHK(ctx, hipMemcpy(f6.data(), s->d_fcomp.p, nf * 6 * 8, hipMemcpyDeviceToHost));
  DRX_HIPCHK(ctx, hipMemcpy(rr.data(), s->d_rout.p, rr.size() * 8, hipMemcpyDeviceToHost));
  const double t_done = now_s();
  std::lock_guard<std::mutex> lk(s->mu);
  for (size_t j = 0; j < nf; ++j) {
    SrvSlot& sl = s->slots[finished[j]];
    SrvDone d;
    d.r.ticket = sl.ticket;
    d.r.K = sl.w.K;
    d.r.status = 0;
    d.r.t_submit = sl.t_submit;
    d.r.t_admit = sl.t_admit;
    d.r.t_done = t_done;
    d.r.out.p_float = f6[6 * j + 0];
    d.r.out.q_float = f6[6 * j + 1];
    d.r.out.log_scale = f6[6 * j + 2];
    d.r.out.delta_float = f6[6 * j + 3];
    d.r.out.log_q = f6[6 * j + 4];
    d.r.out.reserved = 0;
    const int64_t* src = rr.data() + j * 2 * s->kslot;
    d.res.assign(src, src + sl.w.K);
    d.res.insert(d.res.end(), src + s->kslot, src + s->kslot + sl.w.K);
    s->done.push_back(std::move(d));
    sl.used = false;
  }
  return DRX_OK;
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
    std::lock_guard<std::mutex> ctx_lock(s->ctx->mu);
    (void)hipSetDevice(s->ctx->device);
    for (size_t j = 0; j < admitted.size(); ++j) {
      const int rc = admit(s, free_slots[j], admitted[j]);
      if (rc) {
        SrvDone d{};
        d.r.ticket = admitted[j].ticket;
        d.r.status = rc;
        d.r.t_submit = admitted[j].t_submit;
        d.r.t_done = now_s();
        std::lock_guard<std::mutex> lk(s->mu);
        s->done.push_back(std::move(d));
        s->slots[free_slots[j]].used = false;
      }
    }
    int rc = round_once(s, finished);
    if (!rc) rc = harvest(s, finished);
    if (rc) {
      // fail every resident request; keep serving new ones
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
  drx_server* s = new (std::nothrow) drx_server();
  if (!s) return fail(ctx, DRX_ENOMEM, "out of host memory");
  s->ctx = ctx;
  s->cfg = *cfg;
  s->kslot = std::min<int>((int)ctx->primes.size(), kSliceK);  // a served walk takes at most 256 primes
  const int64_t budget = cfg->mem_budget > 0 ? cfg->mem_budget : ((int64_t)1 << 30);
  // resident state + per-round scratch of one slot
  const int64_t per_slot =
      (int64_t)8 * (slot_res(s) * 2 + kFstMax + slot_axr(s) + slot_fdt(s) + kMaxDim * 3 + 6 + 2 * s->kslot) +
      (int64_t)8 * 16 * cfg->round_steps + (int64_t)sizeof(DevWalk) + 2 * sizeof(DevTask) + 3 * sizeof(DevSpan) +
      (int64_t)sizeof(DevRange) * ((cfg->round_steps + kMfRange - 1) / kMfRange);
  int64_t cap = budget / per_slot;
  if (cfg->max_slots > 0) cap = std::min<int64_t>(cap, cfg->max_slots);
  if (cap < 1) {
    delete s;
    return fail(ctx, DRX_ENOMEM, "memory budget too small for one slot");
  }
  s->nslots = (int)cap;
  s->slots.resize(s->nslots);
  s->h_walks.resize(s->nslots);
  const size_t ns = (size_t)s->nslots;
  if (s->d_res.ensure(ns * slot_res(s) * 8) || s->d_fst.ensure(ns * kFstMax * 8) || s->d_axr.ensure(ns * slot_axr(s) * 8) ||
      s->d_fdt.ensure(ns * slot_fdt(s) * 8) ||
      s->d_axi.ensure(ns * kMaxDim * 2 * 8) || s->d_axd.ensure(ns * kMaxDim * 8) || s->d_fout.ensure(ns * 6 * 8) ||
      s->d_rout.ensure(ns * 2 * s->kslot * 8) || s->d_fcomp.ensure(ns * 6 * 8)) {
    delete s;
    return fail(ctx, DRX_ENOMEM, "device allocation for %d slots failed", (int)ns);
  }
  s->th = std::thread(server_loop, s);
  *out = s;
  return DRX_OK;
}

int drx_server_submit(drx_server* s, const drx_walk* w, const int64_t* axis_vals, uint64_t* ticket) {
  if (!s || !w || !ticket) return DRX_EINVAL;
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
    if (w->init_kind == DRX_INIT_GIVEN) return fail(ctx, DRX_EINVAL, "server walks take PCF or identity initial matrices");
    if (w->init_kind == DRX_INIT_PCF && dp.m != 2) return fail(ctx, DRX_EINVAL, "PCF initial matrix needs m=2");
    if (dp.dim > 0 && !axis_vals) return fail(ctx, DRX_EINVAL, "axis values missing");
    rq.axis.assign(axis_vals, axis_vals + 4 * dp.dim);
    rq.epoch = ctx->epoch;
  }
  rq.w = *w;
  rq.t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    rq.ticket = s->next_ticket++;
    *ticket = rq.ticket;
    s->queue.push_back(std::move(rq));
  }
  s->cv.notify_one();
  return DRX_OK;
}

int drx_server_poll(drx_server* s, drx_server_result* res, int64_t* residues, int stride, int max, int* n_done) {
  if (!s || !n_done || max < 0 || (max > 0 && !res)) return DRX_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  int n = 0;
  while (n < max && !s->done.empty()) {
    SrvDone& d = s->done.front();
    if (residues && !d.res.empty()) {
      if (stride < d.r.K) return fail(s->ctx, DRX_EINVAL, "residue stride %d < K=%d", stride, d.r.K);
      int64_t* dst = residues + (size_t)n * 2 * stride;
      std::copy(d.res.begin(), d.res.begin() + d.r.K, dst);
      std::copy(d.res.begin() + d.r.K, d.res.end(), dst + stride);
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
  delete s;
}

}  // extern "C"
