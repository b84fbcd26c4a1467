// action.hip -- the coefficient-accumulating matrix-function action
//
//   out[:, :, k] = sum_{n=1..NC} C[n, k] T_{n-1}(H) V          (k = 1..K)
//
// of chebyshev_action! (src/utils/chebyshev_action.jl:47-125) with its
// fused accumulation broadcast_assign! (src/utils/chebyshev_action.jl:163-178,
// ext/KPMCUDAExt.jl:273-299) and the relative stability guard
// _check_chebyshev_columns (src/utils/chebyshev_iteration.jl:119-132).
//
// The reference adds C[n, k] T_{n-1} V into all K outputs after every step,
// re-reading and re-writing every output block per step (2K block passes per
// step).  Here the recurrence writes into a ring of R (3..16) slots and the
// accumulation runs once per R steps over the whole window: each output word
// is read and written once per window while the R slot words stay in
// registers, so the accumulation costs (R + 2K)/R block passes per step.
// Within a window the per-element sums are added in increasing n, the order
// of the reference's per-step updates.
//
// The guard is evaluated on the device (column norms of the freshly written
// slot at every check_every-th iteration and the last one, against the seed
// norms); the host reads the norms once at the end and reports the first
// failing check exactly as the reference's error() does.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "kpm_internal.h"

namespace kpm {

constexpr int kMaxWindow = 16;

struct WindowSlots {
  const void* s[kMaxWindow];
};

// out_k[row, c] += sum_{j < W} C[n0 + j, k] * slot_j[row, c] for k < K, c < NR.
// One thread per (row, column); slot j in precision PREC (row stride NRpv
// complex), outputs FP64 blocks (row stride NRpo, block stride ostride).
template <int PREC>
__global__ void __launch_bounds__(256)
    k_accum(WindowSlots ws, int W, const double2* __restrict__ C, int64_t NC, int64_t n0, int K,
            double2* __restrict__ out, int64_t ostride, int64_t NH, int NR, int NRpv, int NRpo) {
  const int64_t total = NH * NR;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = idx / NR;
    const int c = (int)(idx - row * NR);
    double vr[kMaxWindow], vi[kMaxWindow];
#pragma unroll
    for (int j = 0; j < kMaxWindow; ++j) {
      vr[j] = vi[j] = 0.0;
      if (j < W) {
        if constexpr (PREC == KPM_VEC_F64) {
          const double2 v = reinterpret_cast<const double2*>(ws.s[j])[row * NRpv + c];
          vr[j] = v.x;
          vi[j] = v.y;
        } else {
          const float2 v = reinterpret_cast<const float2*>(ws.s[j])[row * NRpv + c];
          vr[j] = (double)v.x;
          vi[j] = (double)v.y;
        }
      }
    }
    for (int k = 0; k < K; ++k) {
      double2* o = out + (size_t)k * ostride + row * NRpo + c;
      double2 acc = *o;
      const double2* ck = C + (size_t)k * NC + n0;
#pragma unroll
      for (int j = 0; j < kMaxWindow; ++j) {
        if (j < W) {
          const double2 cc = ck[j];
          // acc += c * v (complex), the reference's out .+= x .* c
          acc.x += cc.x * vr[j] - cc.y * vi[j];
          acc.y += cc.x * vi[j] + cc.y * vr[j];
        }
      }
      *o = acc;
    }
  }
}

static int launch_accum(kpm_ctx* ctx, const WindowSlots& ws, int W, const double2* Cd, int64_t NC,
                        int64_t n0, int K, double2* out, int64_t ostride, int64_t NH, int64_t NR,
                        int prec, int64_t NRpv, int64_t NRpo) {
  const int64_t total = NH * NR;
  const unsigned g = (unsigned)std::min<int64_t>(std::max<int64_t>((total + 255) / 256, 1),
                                                 (int64_t)ctx->num_cus * 32);
  if (prec == KPM_VEC_F32)
    hipLaunchKernelGGL(k_accum<KPM_VEC_F32>, dim3(g), dim3(256), 0, ctx->stream, ws, W, Cd, NC, n0, K,
                       out, ostride, NH, (int)NR, (int)NRpv, (int)NRpo);
  else
    hipLaunchKernelGGL(k_accum<KPM_VEC_F64>, dim3(g), dim3(256), 0, ctx->stream, ws, W, Cd, NC, n0, K,
                       out, ostride, NH, (int)NR, (int)NRpv, (int)NRpo);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

StepConfig tuned_config(kpm_ctx* ctx, const kpm_csr* A, int64_t NR, int prec, int mode);

// Core driver.  seed: the probe block V (read only); out: K FP64 blocks of
// NH x NRpo at stride ostride (zeroed here).  Ring slots come from the
// context's cache.
static int run_action(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t K, const kpm_block* seed,
                      const double* C_host, double2* out, int64_t ostride, int64_t NRpo,
                      int64_t check_every) {
  const int64_t NH = seed->NH, NR = seed->NR, NRpv = seed->NRp;
  const int prec = seed->prec;
  const size_t slot_bytes = (size_t)(NH * NRpv) * elem_bytes(prec);

  // ring size: up to kMaxWindow slots within half of the free memory
  size_t fr = 0, tot = 0;
  KPM_HIP_TRY(hipMemGetInfo(&fr, &tot));
  int64_t R = std::min<int64_t>(kMaxWindow, std::max<int64_t>(NC, 3));
  const int64_t fit = slot_bytes > 0 ? (int64_t)((fr / 2) / slot_bytes) : R;
  R = std::max<int64_t>(3, std::min(R, fit));
  KPM_TRY(ensure(ctx, ctx->slots_l, slot_bytes * (size_t)R));
  auto slot = [&](int64_t m) {  // T_m lives in slot m % R
    return (char*)ctx->slots_l.ptr + (size_t)(m % R) * slot_bytes;
  };
  auto view = [&](int64_t m) {
    kpm_block b;
    b.ctx = ctx;
    b.NH = NH;
    b.NR = NR;
    b.NRp = NRpv;
    b.prec = prec;
    b.d = slot(m);
    b.owned = false;
    return b;
  };

  // coefficient table (NC x K column-major complex) and guard norms
  const int64_t nchk = check_every > 0 ? (NC - 1) / check_every + 3 : 1;
  Buffer cbuf;
  int rc = ensure(ctx, cbuf, (size_t)(NC * K) * sizeof(double2));
  if (rc == KPM_OK) rc = ensure(ctx, ctx->mu, (size_t)(NR * nchk) * sizeof(double2));
  std::vector<int64_t> chk_iter;  // iteration of each guard column (column 0: seed)
  do {
    if (rc != KPM_OK) break;
    if ((rc = hipMemcpyAsync(cbuf.ptr, C_host, (size_t)(NC * K) * sizeof(double2),
                             hipMemcpyHostToDevice, ctx->stream)) != hipSuccess) {
      rc = fail(KPM_EHIP, "coefficient upload failed");
      break;
    }
    if ((rc = hipMemsetAsync(out, 0, (size_t)K * ostride * sizeof(double2), ctx->stream)) != hipSuccess) {
      rc = fail(KPM_EHIP, "output clear failed");
      break;
    }
    if ((rc = hipMemcpyAsync(slot(0), seed->d, slot_bytes, hipMemcpyDeviceToDevice, ctx->stream)) !=
        hipSuccess) {
      rc = fail(KPM_EHIP, "seed copy failed");
      break;
    }
    int G = 0;
    auto norm_into = [&](int64_t m, int64_t iter) -> int {
      kpm_block b = view(m);
      KPM_TRY(launch_dot(ctx, &b, &b, &G));
      Combine c{1, {(int)chk_iter.size(), 0}, {1.0, 0.0}, {-1, -1}, {0.0, 0.0}};
      KPM_TRY(launch_combine(ctx, G, NR, NRpv, c, ctx->mu.ptr, nchk));
      chk_iter.push_back(iter);
      return KPM_OK;
    };
    if (check_every > 0 && (rc = norm_into(0, 0)) != KPM_OK) break;

    int64_t pend = 0;  // first T_m not yet accumulated
    auto flush = [&](int64_t upto) -> int {  // accumulate T_pend .. T_upto
      while (pend <= upto) {
        const int W = (int)std::min<int64_t>(upto - pend + 1, kMaxWindow);
        WindowSlots ws{};
        for (int j = 0; j < W; ++j) ws.s[j] = slot(pend + j);
        KPM_TRY(launch_accum(ctx, ws, W, (const double2*)cbuf.ptr, NC, pend, (int)K, out, ostride,
                             NH, NR, prec, NRpv, NRpo));
        pend += W;
      }
      return KPM_OK;
    };
    const StepConfig cfg = tuned_config(ctx, H, NR, prec, MODE_NONE);
    for (int64_t m = 1; m < NC && rc == KPM_OK; ++m) {
      // T_m overwrites slot m % R, which held T_{m-R}: accumulate it first
      if (m - pend >= R && (rc = flush(m - 1)) != KPM_OK) break;
      kpm_block x = view(m - 1), o = view(m);
      if (m == 1) {
        StepLaunch s{H, &x, nullptr, &o, nullptr, 1.0, 0.0, MODE_NONE};
        rc = launch_step(ctx, s, cfg, nullptr);
      } else {
        kpm_block z = view(m - 2);
        StepLaunch s{H, &x, &z, &o, nullptr, 2.0, -1.0, MODE_NONE};
        rc = launch_step(ctx, s, cfg, nullptr);
      }
      if (rc != KPM_OK) break;
      // guard: iteration m at m == 1 when NC == 2, else every check_every-th and the last
      const bool last = m == NC - 1;
      if (check_every > 0 && ((m == 1 && NC == 2) || (m >= 2 && (m % check_every == 0 || last))))
        rc = norm_into(m, m);
    }
    if (rc != KPM_OK) break;
    if ((rc = flush(NC - 1)) != KPM_OK) break;
    // guard verdict: first failing check, growth relative to the seed norms
    if (check_every > 0) {
      std::vector<double> nr((size_t)(2 * NR * chk_iter.size()));
      if (hipMemcpyAsync(nr.data(), ctx->mu.ptr, nr.size() * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream) != hipSuccess ||
          hipStreamSynchronize(ctx->stream) != hipSuccess) {
        rc = fail(KPM_EHIP, "guard norm download failed");
        break;
      }
      for (size_t q = 1; q < chk_iter.size() && rc == KPM_OK; ++q) {
        double g2 = -INFINITY;
        bool nan = false;
        for (int64_t i = 0; i < NR; ++i) {
          const double ref = nr[2 * i];
          const double v = nr[2 * (i + NR * q)];
          const double r = v / (ref == 0.0 ? 1.0 : ref);
          if (std::isnan(r)) nan = true;
          g2 = std::max(g2, r);
        }
        if (nan || !(g2 <= 1.5 * 1.5)) {
          char msg[512];
          std::snprintf(msg, sizeof msg,
                        "Chebyshev recurrence is unstable at iteration %lld (maximum column growth "
                        "%.17g > 1.5); use a larger eps (rescaling safety margin) in "
                        "rescale/normalizeH, or bound=:gershgorin for BdG operators.",
                        (long long)chk_iter[q], nan ? NAN : std::sqrt(g2));
          rc = fail(KPM_EUNSTABLE, msg);
        }
      }
    }
  } while (false);
  if (cbuf.ptr) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(cbuf.ptr);
  }
  return rc;
}

static int check_action_args(const kpm_csr* H, int64_t NC, int64_t K, int64_t NR, int64_t NH,
                             int64_t check_every) {
  if (NC < 1) return fail(KPM_EINVAL, "C must have at least one coefficient row");
  if (NR < 1) return fail(KPM_EINVAL, "V must have at least one column");
  if (K < 1) return fail(KPM_EINVAL, "C must have at least one coefficient column");
  if (H->n != NH)
    return fail(KPM_EINVAL, "size(Hn, 2) = " + std::to_string(H->n) + " must equal size(V, 1) = " +
                                std::to_string(NH));
  if (check_every < 0) return fail(KPM_EINVAL, "check_every must be >= 0");
  if (NC > INT32_MAX) return fail(KPM_EINVAL, "NC too large");
  return KPM_OK;
}

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_cheb_action(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t K, int64_t NR, int64_t NH,
                    const double* V, const double* C, double* out, int vec_prec,
                    int64_t check_every) {
  if (!ctx || !H || !V || !C || !out) return fail(KPM_EINVAL, "null argument");
  KPM_TRY(check_action_args(H, NC, K, NR, NH, check_every));
  if (vec_prec != KPM_VEC_F64 && vec_prec != KPM_VEC_F32) return fail(KPM_EINVAL, "bad vec_prec");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  kpm_block seed;
  KPM_TRY(work_block(ctx, ctx->work[0], NH, NR, vec_prec, &seed));
  const size_t hbytes = (size_t)(NH * NR) * 16;
  KPM_TRY(ensure(ctx, ctx->staging, hbytes));
  KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, V, hbytes, hipMemcpyHostToDevice, ctx->stream));
  KPM_TRY(launch_upload(ctx, ctx->staging.ptr, &seed));
  const int64_t NRpo = padded_nr(NR, KPM_VEC_F64);
  const int64_t ostride = NH * NRpo;
  Buffer obuf;
  int rc = ensure(ctx, obuf, (size_t)(K * ostride) * sizeof(double2));
  if (rc == KPM_OK)
    rc = run_action(ctx, H, NC, K, &seed, C, (double2*)obuf.ptr, ostride, NRpo, check_every);
  for (int64_t k = 0; k < K && rc == KPM_OK; ++k) {
    kpm_block ob;
    ob.ctx = ctx;
    ob.NH = NH;
    ob.NR = NR;
    ob.NRp = NRpo;
    ob.prec = KPM_VEC_F64;
    ob.d = (double2*)obuf.ptr + k * ostride;
    ob.owned = false;
    if ((rc = launch_download(ctx, &ob, ctx->staging.ptr)) != KPM_OK) break;
    if (hipMemcpyAsync(out + 2 * k * NH * NR, ctx->staging.ptr, hbytes, hipMemcpyDeviceToHost,
                       ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = fail(KPM_EHIP, "output download failed");
  }
  if (obuf.ptr) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(obuf.ptr);
  }
  return rc;
}

int kpm_cheb_action_block(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t K, const kpm_block* V,
                          const double* C, kpm_block* out, int64_t check_every) {
  if (!ctx || !H || !V || !C || !out) return fail(KPM_EINVAL, "null argument");
  KPM_TRY(check_action_args(H, NC, K, V->NR, V->NH, check_every));
  if (out->prec != KPM_VEC_F64 || out->NR != V->NR || out->NH != V->NH * K)
    return fail(KPM_EINVAL, "out must be an FP64 block of (NH*K) x NR");
  if (out->d == V->d) return fail(KPM_EINVAL, "out must not alias V");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  ++ctx->epoch;
  // out stacks the K results: rows [k*NH, (k+1)*NH) hold out[:, :, k]
  return run_action(ctx, H, NC, K, V, C, (double2*)out->d, V->NH * out->NRp, out->NRp, check_every);
}

}  // extern "C"
