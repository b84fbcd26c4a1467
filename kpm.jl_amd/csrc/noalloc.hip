// noalloc.hip -- kpm_1d on device-resident blocks without any allocation.
//
// No counterpart in the reference.  For callers that cannot allocate device
// memory while they run (code under HIP graph capture, tight serving loops):
// the caller passes the probe block, the two recurrence workspaces (the
// reference's alpha_all, src/moment.jl:428-429) and a device moment
// accumulator; the only other device memory the launches use, the
// per-workgroup dot partials, lives in the file-scope array `scratch` below.
// Its device address is looked up once per device when a context is created
// and kept there; kernels take that address, never the array's name.
//
// Every call on a device shares that device's scratch and leaves its
// contents behind: calls that overlap on one device (two streams, two
// contexts, two host threads) race.  Nothing in it has to start at zero:
// each launch writes every partial it later reduces.
#include <mutex>

#include "kpm_internal.h"

namespace kpm {

// two halves of dot partials: workgroups x 2 dots x NRp complex values each
constexpr size_t kScratchBytes = size_t(8) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64] = {};

// Address of `scratch` on ctx->device (the device must be current).  Called
// by kpm_ctx_create, so the lookup happens once per device and never during
// a caller's capture or serving loop.
int noalloc_scratch_init(kpm_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= 64) return fail(KPM_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[ctx->device]) {
    void* p = nullptr;
    KPM_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(KPM_EHIP, "scratch symbol has no device address");
    g_scratch_ptr[ctx->device] = p;
  }
  ctx->scratch = g_scratch_ptr[ctx->device];
  ctx->scratch_bytes = kScratchBytes;
  return KPM_OK;
}

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_1d_block_noalloc(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, const kpm_block* psi,
                         kpm_block* w0, kpm_block* w1, kpm_accum* mu) {
  if (!ctx || !H || !psi || !w0 || !w1 || !mu) return fail(KPM_EINVAL, "null argument");
  if (NC % 2 != 0 || NC < 2) return fail(KPM_EINVAL, "Invalid NC: NC should be even.");
  if (psi->NH != H->n) return fail(KPM_EINVAL, "probe block rows do not match H");
  for (const kpm_block* w : {(const kpm_block*)w0, (const kpm_block*)w1})
    if (w->NH != psi->NH || w->NR != psi->NR || w->prec != psi->prec)
      return fail(KPM_EINVAL, "workspace blocks must have the probe block's shape and precision");
  if (w0->d == w1->d || w0->d == psi->d || w1->d == psi->d)
    return fail(KPM_EINVAL, "probe and workspace blocks must be distinct");
  if (mu->ctx != ctx || mu->NR != psi->NR || mu->NC < NC)
    return fail(KPM_EINVAL, "accumulator must be NR x NC (or wider) on this context");
  if (!ctx->scratch) return fail(KPM_EHIP, "scratch address was not set up for this device");
  // the two partial halves bound the workgroups of a launch
  const size_t per_wg = (size_t)2 * psi->NRp * sizeof(double2);
  const int64_t cap = (int64_t)(ctx->scratch_bytes / 2 / per_wg) / 8 * 8;
  if (cap < 8)
    return fail(KPM_EINVAL, "NR too large for the allocation-free path: the dot partials of 8 "
                            "workgroups (" + std::to_string(16 * per_wg) + " bytes) exceed its " +
                            std::to_string(ctx->scratch_bytes) + "-byte scratch");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  ++ctx->epoch;  // w0, w1 are written
  // w0 <- psi with the copy kernel (no memcpy or memset node under capture)
  const int64_t nwords = psi->NH * psi->NRp * elem_bytes(psi->prec) / 16;
  KPM_TRY(launch_copy16(ctx, psi->d, w0->d, nwords));
  double2* h0 = (double2*)ctx->scratch;
  double2* h1 = (double2*)((char*)ctx->scratch + ctx->scratch_bytes / 2);
  return run_1d_doubling_into(ctx, H, NC, w0, w1, mu->d, h0, h1, (int)std::min<int64_t>(cap, INT32_MAX));
}

}  // extern "C"
