// drx_crt_noalloc: the CRT of drx_crt without any device allocation, for
// callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc). Its only device memory beyond the
// caller's buffers is the staging of residue rows, limb rows and signs, held
// in the file-scope array `scratch` below: a fixed kNoAllocScratch bytes per
// device for the life of the process. Every call on one device shares that
// device's copy, so calls that overlap on one device (two contexts on one
// GPU, or two threads) race; calls on one context are serialised by its
// lock. Both Garner kernels write every limb of a row (zeros up to the row
// stride) and every sign, and the residues are copied in whole, so nothing in
// the scratch has to start at zero.
#include <mutex>
#include <vector>

#include "drx_crt.h"
#include "drx_ctx.h"
#include "drx_kernels.h"

namespace drx {

constexpr size_t kNoAllocScratch = (size_t)8 << 20;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy at its own address). In host code
// the name `scratch` is only a handle for the symbol lookup.
int noalloc_init(drx_ctx* ctx) {
  static std::mutex mu;
  static std::vector<void*> per_device;
  std::lock_guard<std::mutex> lk(mu);
  if ((int)per_device.size() <= ctx->device) per_device.resize(ctx->device + 1, nullptr);
  if (!per_device[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ctx, DRX_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    per_device[ctx->device] = p;
  }
  ctx->scratch = per_device[ctx->device];
  return DRX_OK;
}

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace drx

using namespace drx;

extern "C" int drx_crt_noalloc(drx_ctx* ctx, const int64_t* residues, int n, int K, int stride, uint32_t* limbs,
                               uint8_t* neg) {
  if (!ctx) return DRX_EINVAL;
  if (n < 0 || K < 1 || stride < K || (n > 0 && (!residues || !limbs || !neg)))
    return fail(ctx, DRX_EINVAL, "bad CRT arguments");
  if (n == 0) return DRX_OK;
  const size_t cells = (size_t)n * stride;
  const size_t limb_off = align256(cells * 8), neg_off = align256(limb_off + cells * 4);
  if (neg_off + (size_t)n > kNoAllocScratch)
    return fail(ctx, DRX_ERANGE,
                "n * stride = %zu residues need %zu bytes of staging; the allocation-free CRT holds %zu "
                "(about 699000 residues)", cells, neg_off + (size_t)n, kNoAllocScratch);
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (K > (int)ctx->primes.size()) return fail(ctx, DRX_EINVAL, "K=%d exceeds the %zu primes set", K, ctx->primes.size());
  if (!ctx->scratch) return fail(ctx, DRX_ESTATE, "the allocation-free CRT's scratch was not found on this device");
  // the tables are built when the primes are set; nothing is built here
  const bool wave = K < kCrtMatrixMinK;
  if (wave ? (ctx->kinv < K || !ctx->d_inv.p) : ((int)ctx->garner_w.size() < K || !ctx->d_garner_w.p))
    return fail(ctx, DRX_ESTATE, "CRT tables for K=%d are missing (set the primes first)", K);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < K; ++k) {
      const int64_t r = residues[(size_t)i * stride + k];
      if (r < 0 || r >= (int64_t)ctx->primes[k]) return fail(ctx, DRX_EINVAL, "residue row %d prime %d not canonical", i, k);
    }
  DRX_HIPCHK(ctx, hipSetDevice(ctx->device));
  unsigned char* base = static_cast<unsigned char*>(ctx->scratch);
  hipStream_t s = ctx->stream;
  DRX_HIPCHK(ctx, hipMemcpyAsync(base, residues, cells * 8, hipMemcpyHostToDevice, s));
  GarnerArgs g{};
  g.res = reinterpret_cast<const int64_t*>(base);
  g.res_stride = (uint64_t)stride;
  g.n = n;
  g.K = K;
  g.prime_p = ctx->d_prime_p.as<double>();
  g.prime_pinv = ctx->d_prime_pinv.as<double>();
  g.prime_u = ctx->d_prime_u.as<uint32_t>();
  g.limbs = reinterpret_cast<uint32_t*>(base + limb_off);
  g.neg = base + neg_off;
  if (wave) {
    g.inv_table = ctx->d_inv.as<uint32_t>();
    g.inv_ld = (uint64_t)ctx->kinv;
    DRX_HIPCHK(ctx, launch_garner(g, s));
  } else {
    g.w = ctx->d_garner_w.as<double>();
    DRX_HIPCHK(ctx, launch_garner_wide(g, s));
  }
  DRX_HIPCHK(ctx, hipMemcpyAsync(limbs, base + limb_off, cells * 4, hipMemcpyDeviceToHost, s));
  DRX_HIPCHK(ctx, hipMemcpyAsync(neg, base + neg_off, (size_t)n, hipMemcpyDeviceToHost, s));
  DRX_HIPCHK(ctx, hipStreamSynchronize(s));
  return DRX_OK;
}
