// engine.hip -- the moment engines (src/moment.jl) driven on one HIP stream.
//
//   kpm_1d        kpm_1d! equal-vector doubling      src/moment.jl:420-492
//   kpm_1d_lr     kpm_1d! bra/ket                    src/moment.jl:498-556
//   kpm_1d_current kpm_1d_current!                   src/moment.jl:656-714
//   kpm_2d        kpm_2d! + _kpm_2d_blocked!         src/moment.jl:887-1088
//   kpm_3d        kpm_3d!                            src/moment.jl:1106-1227
//   primitives    chebyshev_iter_single / mul! /     ext/KPMCUDAExt.jl:84-150
//                 columnwise_dot!
//
// Each recurrence step is ONE fused launch (SpMM + the step's moment dots)
// plus one tiny fixed-order reduction launch that writes the moment columns
// with the doubling combination in FP64.  Moments stay on the device until
// one D2H copy at the end (copy_moment_accumulator!, ext/KPMCUDAExt.jl:147).
#include <algorithm>
#include <cstring>
#include <vector>

#include "kpm_internal.h"

namespace kpm {

static int check_block_pair(const kpm_csr* A, const kpm_block* x, const kpm_block* y) {
  if (!A || !x || !y) return fail(KPM_EINVAL, "null argument");
  if (x->NH != A->n || y->NH != A->n) return fail(KPM_EINVAL, "block rows do not match the operator");
  if (x->NR != y->NR || x->prec != y->prec) return fail(KPM_EINVAL, "block shapes or precisions differ");
  return KPM_OK;
}

StepConfig tuned_config(kpm_ctx* ctx, const kpm_csr* A, int64_t NR, int prec, int mode);

static int step(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* x, const kpm_block* z,
                kpm_block* out, const kpm_block* l, double alpha, double beta, int mode, int* grid) {
  StepLaunch s{A, x, z, out, l, alpha, beta, mode};
  return launch_step(ctx, s, tuned_config(ctx, A, x->NR, x->prec, mode), grid);
}

// moments NR x NC on the device, zeroed
static int mu_buffer(kpm_ctx* ctx, int64_t NR, int64_t NC) {
  const size_t bytes = (size_t)(NR * NC) * sizeof(double2);
  KPM_TRY(ensure(ctx, ctx->mu, bytes));
  KPM_HIP_TRY(hipMemsetAsync(ctx->mu.ptr, 0, bytes, ctx->stream));
  return KPM_OK;
}

static int mu_to_host(kpm_ctx* ctx, double* host, int64_t count) {
  KPM_HIP_TRY(hipMemcpyAsync(host, ctx->mu.ptr, (size_t)count * sizeof(double2),
                             hipMemcpyDeviceToHost, ctx->stream));
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return KPM_OK;
}

static int upload_into(kpm_ctx* ctx, const double* host, kpm_block* b) {
  const size_t hbytes = (size_t)(b->NH * b->NR) * 16;
  KPM_TRY(ensure(ctx, ctx->staging, hbytes));
  KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, host, hbytes, hipMemcpyHostToDevice, ctx->stream));
  return launch_upload(ctx, ctx->staging.ptr, b);
}

// Dot steps alternate between two halves of ctx->partials; each launch also
// finishes the previous step's moment columns (PendingCombine), and only the
// last step's columns take a separate combine launch.
struct DotSteps {
  kpm_ctx* ctx;
  const kpm_csr* H;
  int64_t NR, NRp;
  double2* half[2];
  PendingCombine pend;
  bool has_pend = false;
  int k = 0;

  int init(int nd) {
    const size_t per = (size_t)kMaxGridPerCU * std::max(ctx->num_cus, 1) * nd * NRp;
    KPM_TRY(ensure(ctx, ctx->partials, 2 * per * sizeof(double2)));
    half[0] = (double2*)ctx->partials.ptr;
    half[1] = half[0] + per;
    return KPM_OK;
  }
  // one fused pass whose moment columns are described by c
  int step(const kpm_block* x, const kpm_block* z, kpm_block* out, const kpm_block* l, double alpha,
           double beta, int mode, const Combine& c) {
    StepLaunch s{H, x, z, out, l, alpha, beta, mode};
    s.part_out = half[k & 1];
    s.pend = has_pend ? &pend : nullptr;
    int G = 0;
    KPM_TRY(launch_step(ctx, s, tuned_config(ctx, H, x->NR, x->prec, mode), &G));
    pend.part = half[k & 1];
    pend.mu = (double2*)ctx->mu.ptr;
    pend.cb = c;
    pend.G = G;
    pend.NRp = (int)NRp;
    pend.NR = (int)NR;
    has_pend = true;
    ++k;
    return KPM_OK;
  }
  int finish() {
    if (!has_pend) return KPM_OK;
    has_pend = false;
    return launch_combine_from(ctx, pend.part, pend.G, NR, NRp, pend.cb, ctx->mu.ptr);
  }
};

// ---- kpm_1d! equal-vector, on a device probe block -------------------------
// w0 holds psi on entry; w1 is scratch.  mu (device) receives NR x NC.
static int run_1d_doubling(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, kpm_block* w0,
                           kpm_block* w1) {
  const int64_t NR = w0->NR, NRp = w0->NRp;
  KPM_TRY(mu_buffer(ctx, NR, NC));
  DotSteps ds{ctx, H, NR, NRp, {nullptr, nullptr}};
  KPM_TRY(ds.init(2));
  kpm_block* v[2] = {w0, w1};
  // alpha_1 = H alpha_0 with mu0 = <a0|a0>, mu1 = <a0|a1> (:442-449)
  Combine c0{2, {0, 1}, {1.0, 1.0}, {-1, -1}, {0.0, 0.0}};
  KPM_TRY(ds.step(v[0], nullptr, v[1], nullptr, 1.0, 0.0, MODE_DBL, c0));
  int ip = 1, ipp = 0;
  for (int64_t n = 2; n <= NC / 2; ++n) {  // :460-487
    // mu_{2n-2} = 2<a_{n-1}|a_{n-1}> - mu0 ; mu_{2n-1} = 2<a_{n-1}|a_n> - mu1
    Combine c{2, {(int)(2 * n - 2), (int)(2 * n - 1)}, {2.0, 2.0}, {0, 1}, {-1.0, -1.0}};
    KPM_TRY(ds.step(v[ip], v[ipp], v[ipp], nullptr, 2.0, -1.0, MODE_DBL, c));
    ip = 1 - ip;
    ipp = 1 - ipp;
  }
  return ds.finish();
}

// ---- bra/ket and current: one dot per step against l ------------------------
static int run_1d_lr(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, const kpm_block* l,
                     kpm_block* w0, kpm_block* w1) {
  const int64_t NR = w0->NR, NRp = w0->NRp;
  KPM_TRY(mu_buffer(ctx, NR, NC));
  int G = 0;
  KPM_TRY(launch_dot(ctx, l, w0, &G));  // mu0 = <l|a0>
  Combine c0{1, {0, 0}, {1.0, 0.0}, {-1, -1}, {0.0, 0.0}};
  KPM_TRY(launch_combine(ctx, G, NR, NRp, c0, ctx->mu.ptr, NC));
  DotSteps ds{ctx, H, NR, NRp, {nullptr, nullptr}};
  KPM_TRY(ds.init(1));
  kpm_block* v[2] = {w0, w1};
  Combine c1{1, {1, 0}, {1.0, 0.0}, {-1, -1}, {0.0, 0.0}};
  KPM_TRY(ds.step(v[0], nullptr, v[1], l, 1.0, 0.0, MODE_LR, c1));  // a1 = H a0, mu1
  int ip = 1, ipp = 0;
  for (int64_t n = 3; n <= NC; ++n) {
    Combine c{1, {(int)(n - 1), 0}, {1.0, 0.0}, {-1, -1}, {0.0, 0.0}};
    KPM_TRY(ds.step(v[ip], v[ipp], v[ipp], l, 2.0, -1.0, MODE_LR, c));
    ip = 1 - ip;
    ipp = 1 - ipp;
  }
  return ds.finish();
}

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_1d(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t NR, int64_t NH, const double* psi,
           double* mu_all, int vec_prec) {
  if (!ctx || !H || !psi || !mu_all) return fail(KPM_EINVAL, "null argument");
  if (NC % 2 != 0 || NC < 2) return fail(KPM_EINVAL, "Invalid NC: NC should be even.");
  if (NR <= 0 || NH != H->n) return fail(KPM_EINVAL, "Invalid `psi_in` size: expecting (NH, NR)");
  if (vec_prec != KPM_VEC_F64 && vec_prec != KPM_VEC_F32) return fail(KPM_EINVAL, "bad vec_prec");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  kpm_block w0, w1;
  KPM_TRY(work_block(ctx, ctx->work[0], NH, NR, vec_prec, &w0));
  KPM_TRY(work_block(ctx, ctx->work[1], NH, NR, vec_prec, &w1));
  KPM_TRY(upload_into(ctx, psi, &w0));
  KPM_TRY(run_1d_doubling(ctx, H, NC, &w0, &w1));
  return mu_to_host(ctx, mu_all, NR * NC);
}

int kpm_1d_block(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, const kpm_block* psi, double* mu_all) {
  if (!ctx || !H || !psi || !mu_all) return fail(KPM_EINVAL, "null argument");
  if (NC % 2 != 0 || NC < 2) return fail(KPM_EINVAL, "Invalid NC: NC should be even.");
  if (psi->NH != H->n) return fail(KPM_EINVAL, "probe block rows do not match H");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  kpm_block w0, w1;
  KPM_TRY(work_block(ctx, ctx->work[0], psi->NH, psi->NR, psi->prec, &w0));
  KPM_TRY(work_block(ctx, ctx->work[1], psi->NH, psi->NR, psi->prec, &w1));
  KPM_TRY(kpm_block_copy(&w0, psi));
  KPM_TRY(run_1d_doubling(ctx, H, NC, &w0, &w1));
  return mu_to_host(ctx, mu_all, psi->NR * NC);
}

int kpm_1d_lr(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t NR, int64_t NH,
              const double* psi_l, const double* psi_r, double* mu_all, int vec_prec) {
  if (!ctx || !H || !psi_l || !psi_r || !mu_all) return fail(KPM_EINVAL, "null argument");
  if (NC < 2) return fail(KPM_EINVAL, "Invalid NC: NC should be at least 2.");
  if (NR <= 0 || NH != H->n) return fail(KPM_EINVAL, "Invalid probe size: expecting (NH, NR)");
  if (vec_prec != KPM_VEC_F64 && vec_prec != KPM_VEC_F32) return fail(KPM_EINVAL, "bad vec_prec");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  kpm_block w0, w1, l;
  KPM_TRY(work_block(ctx, ctx->work[0], NH, NR, vec_prec, &w0));
  KPM_TRY(work_block(ctx, ctx->work[1], NH, NR, vec_prec, &w1));
  KPM_TRY(work_block(ctx, ctx->lwork, NH, NR, vec_prec, &l));
  KPM_TRY(upload_into(ctx, psi_l, &l));
  KPM_TRY(upload_into(ctx, psi_r, &w0));
  KPM_TRY(run_1d_lr(ctx, H, NC, &l, &w0, &w1));
  return mu_to_host(ctx, mu_all, NR * NC);
}

int kpm_1d_lr_block(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, const kpm_block* psi_l,
                    const kpm_block* psi_r, double* mu_all) {
  if (!ctx || !H || !psi_l || !psi_r || !mu_all) return fail(KPM_EINVAL, "null argument");
  if (NC < 2) return fail(KPM_EINVAL, "Invalid NC: NC should be at least 2.");
  KPM_TRY(check_block_pair(H, psi_l, psi_r));
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  kpm_block w0, w1;
  KPM_TRY(work_block(ctx, ctx->work[0], psi_r->NH, psi_r->NR, psi_r->prec, &w0));
  KPM_TRY(work_block(ctx, ctx->work[1], psi_r->NH, psi_r->NR, psi_r->prec, &w1));
  KPM_TRY(kpm_block_copy(&w0, psi_r));
  KPM_TRY(run_1d_lr(ctx, H, NC, psi_l, &w0, &w1));
  return mu_to_host(ctx, mu_all, psi_r->NR * NC);
}

int kpm_1d_current(kpm_ctx* ctx, const kpm_csr* H, const kpm_csr* Jadj, int64_t NC, int64_t NR,
                   int64_t NH, const double* psi, double* mu_all, int vec_prec) {
  if (!ctx || !H || !Jadj || !psi || !mu_all) return fail(KPM_EINVAL, "null argument");
  if (NC < 2) return fail(KPM_EINVAL, "Invalid NC: NC should be at least 2.");
  if (NR <= 0 || NH != H->n || Jadj->n != NH) return fail(KPM_EINVAL, "Invalid `psi_in` size");
  if (vec_prec != KPM_VEC_F64 && vec_prec != KPM_VEC_F32) return fail(KPM_EINVAL, "bad vec_prec");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  kpm_block w0, w1, l;
  KPM_TRY(work_block(ctx, ctx->work[0], NH, NR, vec_prec, &w0));
  KPM_TRY(work_block(ctx, ctx->work[1], NH, NR, vec_prec, &w1));
  KPM_TRY(work_block(ctx, ctx->lwork, NH, NR, vec_prec, &l));
  KPM_TRY(upload_into(ctx, psi, &w0));
  // J_alpha^dagger psi once (:680), then the bra/ket recurrence against it
  KPM_TRY(step(ctx, Jadj, &w0, nullptr, &l, nullptr, 1.0, 0.0, MODE_NONE, nullptr));
  KPM_TRY(run_1d_lr(ctx, H, NC, &l, &w0, &w1));
  return mu_to_host(ctx, mu_all, NR * NC);
}

// ---- primitives -----------------------------------------------------------------

int kpm_cheb_step(kpm_ctx* ctx, const kpm_csr* H, const kpm_block* Vpp, const kpm_block* Vp,
                  kpm_block* Vout) {
  if (!ctx) return fail(KPM_EINVAL, "null context");
  KPM_TRY(check_block_pair(H, Vpp, Vp));
  KPM_TRY(check_block_pair(H, Vp, Vout));
  if (Vp->d == Vout->d) return fail(KPM_EINVAL, "V_p must not alias V_out");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  return step(ctx, H, Vp, Vpp, Vout, nullptr, 2.0, -1.0, MODE_NONE, nullptr);
}

int kpm_spmm(kpm_ctx* ctx, const kpm_csr* A, double alpha, const kpm_block* X, double beta,
             kpm_block* Y) {
  if (!ctx) return fail(KPM_EINVAL, "null context");
  KPM_TRY(check_block_pair(A, X, Y));
  if (X->d == Y->d) return fail(KPM_EINVAL, "X must not alias Y");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  return step(ctx, A, X, beta != 0.0 ? Y : nullptr, Y, nullptr, alpha, beta, MODE_NONE, nullptr);
}

int kpm_colwise_dot(kpm_ctx* ctx, double* target, int64_t NC, int64_t n, const kpm_block* A,
                    const kpm_block* B, double alpha, int64_t beta_col, double beta_scale) {
  if (!ctx || !target || !A || !B) return fail(KPM_EINVAL, "null argument");
  if (A->NH != B->NH || A->NR != B->NR || A->prec != B->prec)
    return fail(KPM_EINVAL, "block shapes or precisions differ");
  if (n < 1 || n > NC || beta_col < 0 || beta_col > NC) return fail(KPM_EINVAL, "column out of range");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int64_t NR = A->NR;
  KPM_TRY(ensure(ctx, ctx->mu, (size_t)NR * sizeof(double2)));
  int G = 0;
  KPM_TRY(launch_dot(ctx, A, B, &G));
  Combine c{1, {0, 0}, {1.0, 0.0}, {-1, -1}, {0.0, 0.0}};
  KPM_TRY(launch_combine(ctx, G, NR, A->NRp, c, ctx->mu.ptr, 1));
  std::vector<double> d(2 * NR);
  KPM_TRY(mu_to_host(ctx, d.data(), NR));
  for (int64_t i = 0; i < NR; ++i) {
    double br = 0.0, bi = 0.0;
    if (beta_col != 0) {
      br = beta_scale * target[2 * (i + NR * (beta_col - 1))];
      bi = beta_scale * target[2 * (i + NR * (beta_col - 1)) + 1];
    }
    target[2 * (i + NR * (n - 1))] = alpha * d[2 * i] + br;
    target[2 * (i + NR * (n - 1)) + 1] = alpha * d[2 * i + 1] + bi;
  }
  return KPM_OK;
}

int kpm_bench_step(kpm_ctx* ctx, const kpm_csr* H, kpm_block* X, kpm_block* Z, int64_t nsteps,
                   float* ms) {
  if (!ctx || !ms || nsteps < 1) return fail(KPM_EINVAL, "bad arguments");
  KPM_TRY(check_block_pair(H, X, Z));
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  KPM_HIP_TRY(hipEventRecord(ctx->events[14], ctx->stream));
  for (int64_t s = 0; s < nsteps; ++s)
    KPM_TRY(step(ctx, H, X, Z, Z, nullptr, 2.0, -1.0, MODE_DBL, nullptr));
  KPM_HIP_TRY(hipEventRecord(ctx->events[15], ctx->stream));
  KPM_HIP_TRY(hipEventSynchronize(ctx->events[15]));
  KPM_HIP_TRY(hipEventElapsedTime(ms, ctx->events[14], ctx->events[15]));
  return KPM_OK;
}

int kpm_bench_copy(kpm_ctx* ctx, kpm_block* dst, const kpm_block* src, int64_t reps, float* ms) {
  if (!ctx || !dst || !src || !ms || reps < 1) return fail(KPM_EINVAL, "bad arguments");
  if (dst->NH != src->NH || dst->NRp != src->NRp || dst->prec != src->prec)
    return fail(KPM_EINVAL, "block shapes differ");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int64_t nwords = dst->NH * dst->NRp * elem_bytes(dst->prec) / 16;
  KPM_HIP_TRY(hipEventRecord(ctx->events[14], ctx->stream));
  for (int64_t r = 0; r < reps; ++r) KPM_TRY(launch_copy16(ctx, src->d, dst->d, nwords));
  KPM_HIP_TRY(hipEventRecord(ctx->events[15], ctx->stream));
  KPM_HIP_TRY(hipEventSynchronize(ctx->events[15]));
  KPM_HIP_TRY(hipEventElapsedTime(ms, ctx->events[14], ctx->events[15]));
  return KPM_OK;
}

}  // extern "C"
