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

#include "dot_steps.h"

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
  ++ctx->counters[KPM_COUNT_SYNC];
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return KPM_OK;
}

static int upload_into(kpm_ctx* ctx, const double* host, kpm_block* b) {
  const size_t hbytes = (size_t)(b->NH * b->NR) * 16;
  KPM_TRY(ensure(ctx, ctx->staging, hbytes));
  KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, host, hbytes, hipMemcpyHostToDevice, ctx->stream));
  return launch_upload(ctx, ctx->staging.ptr, b);
}

int time_engine_steps(kpm_ctx* ctx, const kpm_csr* H, kpm_block* X, kpm_block* Z,
                      const StepConfig& cfg, int reps, float* ms, int mode) {
  const bool avg = mode == MODE_DBLA;
  KPM_TRY(ensure(ctx, ctx->mu, (size_t)(X->NR * 2) * sizeof(double2)));
  DotSteps ds{ctx, H, avg ? 1 : X->NR, avg ? 1 : X->NRp, {nullptr, nullptr}};
  ds.cfg = &cfg;
  KPM_TRY(ds.init(2, X, mode));
  const Combine c{2, {0, 1}, {2.0, 2.0}, {-1, -1}, {0.0, 0.0}};
  for (int r = 0; r < 2; ++r) KPM_TRY(ds.step(X, Z, Z, nullptr, 2.0, -1.0, mode, c));
  KPM_HIP_TRY(hipEventRecord(ctx->events[12], ctx->stream));
  for (int r = 0; r < reps; ++r) KPM_TRY(ds.step(X, Z, Z, nullptr, 2.0, -1.0, mode, c));
  KPM_TRY(ds.finish());
  KPM_HIP_TRY(hipEventRecord(ctx->events[13], ctx->stream));
  KPM_HIP_TRY(hipEventSynchronize(ctx->events[13]));
  KPM_HIP_TRY(hipEventElapsedTime(ms, ctx->events[12], ctx->events[13]));
  return KPM_OK;
}

// ---- kpm_1d! equal-vector, on a device probe block -------------------------
// w0 holds psi on entry; w1 is scratch.  mu_dev receives NR x NC.
}  // namespace kpm

int kpm::run_1d_doubling_into(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, kpm_block* w0,
                              kpm_block* w1, void* mu_dev, double2* part0, double2* part1,
                              int max_grid) {
  const int64_t NR = w0->NR, NRp = w0->NRp;
  DotSteps ds{ctx, H, NR, NRp, {nullptr, nullptr}};
  if (part0 && part1) ds.init_external(part0, part1, max_grid);
  else KPM_TRY(ds.init(2, w0, MODE_DBL));
  ds.mu_dev = (double2*)mu_dev;
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

namespace kpm {

static int run_1d_doubling(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, kpm_block* w0,
                           kpm_block* w1) {
  KPM_TRY(mu_buffer(ctx, w0->NR, NC));
  return run_1d_doubling_into(ctx, H, NC, w0, w1, ctx->mu.ptr, nullptr, nullptr, 0);
}

// The doubling recurrence of run_1d_doubling_into with every dot summed over
// the probe columns in the pass itself (MODE_DBLA): ctx->mu receives the
// 1 x NC probe sums sum_r mu[r, n], the numerator of kpm_1d's avg_output
// (src/moment.jl:410-416).  The doubling combination is linear, so it holds
// for the probe sums as for each probe.
static int run_1d_doubling_avg(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, kpm_block* w0,
                               kpm_block* w1) {
  KPM_TRY(mu_buffer(ctx, 1, NC));
  DotSteps ds{ctx, H, 1, 1, {nullptr, nullptr}};
  KPM_TRY(ds.init(2, w0, MODE_DBLA));
  kpm_block* v[2] = {w0, w1};
  Combine c0{2, {0, 1}, {1.0, 1.0}, {-1, -1}, {0.0, 0.0}};
  KPM_TRY(ds.step(v[0], nullptr, v[1], nullptr, 1.0, 0.0, MODE_DBLA, c0));
  int ip = 1, ipp = 0;
  for (int64_t n = 2; n <= NC / 2; ++n) {
    Combine c{2, {(int)(2 * n - 2), (int)(2 * n - 1)}, {2.0, 2.0}, {0, 1}, {-1.0, -1.0}};
    KPM_TRY(ds.step(v[ip], v[ipp], v[ipp], nullptr, 2.0, -1.0, MODE_DBLA, c));
    ip = 1 - ip;
    ipp = 1 - ipp;
  }
  return ds.finish();
}

// probe sums -> probe averages (src/moment.jl:411: sum(mu_all, dims = 1) ./ NR)
static int avg_to_host(kpm_ctx* ctx, double* host, int64_t NC, int64_t NR) {
  KPM_TRY(mu_to_host(ctx, host, NC));
  const double inv = 1.0 / (double)NR;
  for (int64_t i = 0; i < 2 * NC; ++i) host[i] *= inv;
  return KPM_OK;
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
  KPM_TRY(ds.init(1, w0, MODE_LR));
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

int kpm_1d_avg(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t NR, int64_t NH,
               const double* psi, double* mu_avg, int vec_prec) {
  if (!ctx || !H || !psi || !mu_avg) return fail(KPM_EINVAL, "null argument");
  if (NC % 2 != 0 || NC < 2) return fail(KPM_EINVAL, "Invalid NC: NC should be even.");
  if (NR <= 0 || NH != H->n) return fail(KPM_EINVAL, "Invalid `psi_in` size: expecting (NH, NR)");
  if (vec_prec != KPM_VEC_F64 && vec_prec != KPM_VEC_F32) return fail(KPM_EINVAL, "bad vec_prec");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  kpm_block w0, w1;
  KPM_TRY(work_block(ctx, ctx->work[0], NH, NR, vec_prec, &w0));
  KPM_TRY(work_block(ctx, ctx->work[1], NH, NR, vec_prec, &w1));
  KPM_TRY(upload_into(ctx, psi, &w0));
  KPM_TRY(run_1d_doubling_avg(ctx, H, NC, &w0, &w1));
  return avg_to_host(ctx, mu_avg, NC, NR);
}

int kpm_1d_block_avg(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, const kpm_block* psi,
                     double* mu_avg) {
  if (!ctx || !H || !psi || !mu_avg) return fail(KPM_EINVAL, "null argument");
  if (NC % 2 != 0 || NC < 2) return fail(KPM_EINVAL, "Invalid NC: NC should be even.");
  if (psi->NH != H->n) return fail(KPM_EINVAL, "probe block rows do not match H");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  kpm_block w0, w1;
  KPM_TRY(work_block(ctx, ctx->work[0], psi->NH, psi->NR, psi->prec, &w0));
  KPM_TRY(work_block(ctx, ctx->work[1], psi->NH, psi->NR, psi->prec, &w1));
  KPM_TRY(kpm_block_copy(&w0, psi));
  KPM_TRY(run_1d_doubling_avg(ctx, H, NC, &w0, &w1));
  return avg_to_host(ctx, mu_avg, NC, psi->NR);
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
// The reference's device loop (src/moment.jl:440-489 on ext/KPMCUDAExt.jl:84-150)
// calls chebyshev_iter_single and then columnwise_dot! twice per step.  Here
// every primitive step is the fused pass that also reduces <Vp|Vp> and
// <Vp|Vout>; the following columnwise_dot! calls on those blocks combine the
// cached partials (O(grid * NR)) instead of re-reading the blocks, and the
// combine runs inside the next primitive step's launch.

}  // extern "C"

namespace kpm {

// run the combine waiting for the next primitive step as a launch of its own
static int prim_flush(kpm_ctx* ctx) {
  kpm_prim_state& P = ctx->prim;
  if (!P.pending) return KPM_OK;
  P.pending = false;
  return launch_combine_from(ctx, (const double2*)P.pend_part, P.pend_G, P.pend_NR, P.pend_NRp,
                             P.pend_cb, P.pend_mu);
}

// out = alpha*A*x + beta*z (z may be null or alias out) with d0 = <x|x>,
// d1 = <x|out> into the next half of the primitive partials, running the
// pending combine in the same launch
static int prim_step(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* x, const kpm_block* z,
                     kpm_block* out, double alpha, double beta) {
  kpm_prim_state& P = ctx->prim;
  const int cap = partials_grid(ctx, A, x->NH, x->NR, x->prec, MODE_DBL);
  const size_t half_bytes = (size_t)cap * 2 * x->NRp * sizeof(double2);
  if (P.part.bytes < 2 * half_bytes) {  // growing frees the halves the pending combine reads
    KPM_TRY(prim_flush(ctx));
    P.cached = false;
    KPM_TRY(ensure(ctx, P.part, 2 * half_bytes));
  }
  const int h = P.k & 1;
  double2* part = (double2*)((char*)P.part.ptr + h * (P.part.bytes / 2));
  StepLaunch s{A, x, z, out, nullptr, alpha, beta, MODE_DBL};
  s.part_out = part;
  s.max_grid = (int)(P.part.bytes / 2 / (2 * x->NRp * sizeof(double2)));
  PendingCombine pc;
  if (P.pending) {
    pc.part = (const double2*)P.pend_part;
    pc.mu = (double2*)P.pend_mu;
    pc.cb = P.pend_cb;
    pc.G = P.pend_G;
    pc.NRp = P.pend_NRp;
    pc.NR = P.pend_NR;
    s.pend = &pc;
  }
  int G = 0;
  KPM_TRY(launch_step(ctx, s, tuned_config(ctx, A, x->NR, x->prec, MODE_DBL), &G));
  P.pending = false;
  ++ctx->epoch;  // out changed: dots cached before this step are stale
  P.cached = true;
  P.vp = x->d;
  P.vout = out->d;
  P.epoch = ctx->epoch;
  P.part_ptr = part;
  P.G = G;
  P.NRp = (int)x->NRp;
  P.NR = (int)x->NR;
  ++P.k;
  return KPM_OK;
}

// which cached dot of the last primitive step <A|B> is (-1: none)
static int prim_lookup(kpm_ctx* ctx, const kpm_block* A, const kpm_block* B) {
  const kpm_prim_state& P = ctx->prim;
  if (!P.cached || P.epoch != ctx->epoch || A->d != P.vp || A->NR != P.NR || A->NRp != P.NRp)
    return -1;
  if (B->d == P.vp) return 0;
  if (B->d == P.vout) return 1;
  return -1;
}

static int check_dot_args(const kpm_block* A, const kpm_block* B, int64_t NC, int64_t n,
                          int64_t beta_col) {
  if (A->NH != B->NH || A->NR != B->NR || A->prec != B->prec)
    return fail(KPM_EINVAL, "block shapes or precisions differ");
  if (n < 1 || n > NC || beta_col < 0 || beta_col > NC) return fail(KPM_EINVAL, "column out of range");
  return KPM_OK;
}

}  // namespace kpm



extern "C" {

int kpm_cheb_step(kpm_ctx* ctx, const kpm_csr* H, const kpm_block* Vpp, const kpm_block* Vp,
                  kpm_block* Vout) {
  if (!ctx) return fail(KPM_EINVAL, "null context");
  KPM_TRY(check_block_pair(H, Vpp, Vp));
  KPM_TRY(check_block_pair(H, Vp, Vout));
  if (Vp->d == Vout->d) return fail(KPM_EINVAL, "V_p must not alias V_out");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  return prim_step(ctx, H, Vp, Vpp, Vout, 2.0, -1.0);
}

int kpm_spmm(kpm_ctx* ctx, const kpm_csr* A, double alpha, const kpm_block* X, double beta,
             kpm_block* Y) {
  if (!ctx) return fail(KPM_EINVAL, "null context");
  KPM_TRY(check_block_pair(A, X, Y));
  if (X->d == Y->d) return fail(KPM_EINVAL, "X must not alias Y");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  return prim_step(ctx, A, X, beta != 0.0 ? Y : nullptr, Y, alpha, beta);
}

int kpm_colwise_dot(kpm_ctx* ctx, double* target, int64_t NC, int64_t n, const kpm_block* A,
                    const kpm_block* B, double alpha, int64_t beta_col, double beta_scale) {
  if (!ctx || !target || !A || !B) return fail(KPM_EINVAL, "null argument");
  KPM_TRY(check_dot_args(A, B, NC, n, beta_col));
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int64_t NR = A->NR;
  KPM_TRY(ensure(ctx, ctx->mu, (size_t)NR * sizeof(double2)));
  const Combine c{1, {0, 0}, {1.0, 0.0}, {-1, -1}, {0.0, 0.0}};
  const int d = prim_lookup(ctx, A, B);
  if (d >= 0) {  // the last primitive step reduced this dot already
    const kpm_prim_state& P = ctx->prim;
    // one dot out of the step's pair: entry 0 reads dot d of stride 2
    Combine cd{2, {0, 0}, {0.0, 0.0}, {-1, -1}, {0.0, 0.0}};
    cd.tcol[d] = 0;
    cd.alpha[d] = 1.0;
    cd.tcol[1 - d] = 1;  // scratch column
    KPM_TRY(ensure(ctx, ctx->mu, (size_t)2 * NR * sizeof(double2)));
    KPM_TRY(launch_combine_from(ctx, (const double2*)P.part_ptr, P.G, NR, P.NRp, cd, ctx->mu.ptr));
  } else {
    int G = 0;
    KPM_TRY(launch_dot(ctx, A, B, &G));
    KPM_TRY(launch_combine(ctx, G, NR, A->NRp, c, ctx->mu.ptr, 1));
  }
  std::vector<double> dv(2 * NR);
  KPM_TRY(mu_to_host(ctx, dv.data(), NR));
  for (int64_t i = 0; i < NR; ++i) {
    double br = 0.0, bi = 0.0;
    if (beta_col != 0) {
      br = beta_scale * target[2 * (i + NR * (beta_col - 1))];
      bi = beta_scale * target[2 * (i + NR * (beta_col - 1)) + 1];
    }
    target[2 * (i + NR * (n - 1))] = alpha * dv[2 * i] + br;
    target[2 * (i + NR * (n - 1)) + 1] = alpha * dv[2 * i + 1] + bi;
  }
  return KPM_OK;
}

int kpm_accum_create(kpm_ctx* ctx, int64_t NR, int64_t NC, kpm_accum** out) {
  if (!ctx || !out) return fail(KPM_EINVAL, "null argument");
  if (NR <= 0 || NC <= 0) return fail(KPM_EINVAL, "accumulator dimensions must be positive");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const size_t bytes = (size_t)(NR * (NC + 1)) * sizeof(double2);
  kpm_accum* a = new kpm_accum();
  a->ctx = ctx;
  a->NR = NR;
  a->NC = NC;
  if (hipMalloc(&a->d, bytes) != hipSuccess) {
    (void)hipGetLastError();
    delete a;
    return fail(KPM_ENOMEM, "hipMalloc of the moment accumulator failed");
  }
  hipError_t e = hipMemsetAsync(a->d, 0, bytes, ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(a->d);
    delete a;
    return fail(KPM_EHIP, std::string("accumulator clear failed: ") + hipGetErrorString(e));
  }
  *out = a;
  return KPM_OK;
}

int kpm_accum_destroy(kpm_accum* a) {
  if (!a) return KPM_OK;
  kpm_ctx* ctx = a->ctx;
  (void)hipSetDevice(ctx->device);
  if (ctx->prim.pending && ctx->prim.pend_mu == a->d) ctx->prim.pending = false;  // nobody reads it
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(a->d);
  delete a;
  return KPM_OK;
}

int kpm_accum_dot(kpm_ctx* ctx, kpm_accum* acc, int64_t n, const kpm_block* A, const kpm_block* B,
                  double alpha, int64_t beta_col, double beta_scale) {
  if (!ctx || !acc || !A || !B) return fail(KPM_EINVAL, "null argument");
  if (acc->ctx != ctx) return fail(KPM_EINVAL, "accumulator belongs to another context");
  KPM_TRY(check_dot_args(A, B, acc->NC, n, beta_col));
  if (A->NR != acc->NR) return fail(KPM_EINVAL, "accumulator rows must equal the probe columns");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  kpm_prim_state& P = ctx->prim;
  const int tcol = (int)(n - 1), bcol = (int)(beta_col - 1);
  const int d = prim_lookup(ctx, A, B);
  if (d >= 0) {
    // join the waiting combine when it reduces the same partials into the
    // same accumulator and no entry reads a column another entry writes
    if (P.pending) {
      const int o = 1 - d;
      const bool join = P.pend_mu == acc->d && P.pend_part == P.part_ptr && !P.pend_used[d] &&
                        !(P.pend_used[o] && (P.pend_cb.tcol[o] == tcol || P.pend_cb.tcol[o] == bcol ||
                                             P.pend_cb.bcol[o] == tcol));
      if (!join) KPM_TRY(prim_flush(ctx));
    }
    if (!P.pending) {
      const int scratch = (int)acc->NC;
      P.pending = true;
      P.pend_mu = acc->d;
      P.pend_part = P.part_ptr;
      P.pend_G = P.G;
      P.pend_NRp = P.NRp;
      P.pend_NR = P.NR;
      P.pend_cb = Combine{2, {scratch, scratch}, {0.0, 0.0}, {-1, -1}, {0.0, 0.0}};
      P.pend_used[0] = P.pend_used[1] = false;
    }
    P.pend_cb.tcol[d] = tcol;
    P.pend_cb.alpha[d] = alpha;
    P.pend_cb.bcol[d] = bcol;
    P.pend_cb.bscale[d] = beta_scale;
    P.pend_used[d] = true;
    return KPM_OK;
  }
  // not a dot of the last step: one dot-only pass, reduced on the device
  KPM_TRY(prim_flush(ctx));
  int G = 0;
  KPM_TRY(launch_dot(ctx, A, B, &G));
  const Combine c{1, {tcol, 0}, {alpha, 0.0}, {bcol, -1}, {beta_scale, 0.0}};
  return launch_combine(ctx, G, A->NR, A->NRp, c, acc->d, acc->NC);
}

int kpm_accum_download(kpm_ctx* ctx, const kpm_accum* acc, double* host) {
  if (!ctx || !acc || !host) return fail(KPM_EINVAL, "null argument");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  if (ctx->prim.pending && ctx->prim.pend_mu == acc->d) KPM_TRY(prim_flush(ctx));
  KPM_HIP_TRY(hipMemcpyAsync(host, acc->d, (size_t)(acc->NR * acc->NC) * sizeof(double2),
                             hipMemcpyDeviceToHost, ctx->stream));
  ++ctx->counters[KPM_COUNT_SYNC];
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return KPM_OK;
}

int kpm_bench_step(kpm_ctx* ctx, const kpm_csr* H, kpm_block* X, kpm_block* Z, int64_t nsteps,
                   float* ms) {
  return kpm_bench_step_mode(ctx, H, X, Z, nsteps, 0, ms);
}

int kpm_bench_step_mode(kpm_ctx* ctx, const kpm_csr* H, kpm_block* X, kpm_block* Z,
                        int64_t nsteps, int mode, float* ms) {
  if (!ctx || !ms || nsteps < 1 || mode < 0 || mode > 1) return fail(KPM_EINVAL, "bad arguments");
  KPM_TRY(check_block_pair(H, X, Z));
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  ++ctx->epoch;
  const int m = mode ? MODE_DBLA : MODE_DBL;
  KPM_HIP_TRY(hipEventRecord(ctx->events[14], ctx->stream));
  for (int64_t s = 0; s < nsteps; ++s)
    KPM_TRY(step(ctx, H, X, Z, Z, nullptr, 2.0, -1.0, m, nullptr));
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
  ++ctx->epoch;
  const int64_t nwords = dst->NH * dst->NRp * elem_bytes(dst->prec) / 16;
  KPM_HIP_TRY(hipEventRecord(ctx->events[14], ctx->stream));
  for (int64_t r = 0; r < reps; ++r) KPM_TRY(launch_copy16(ctx, src->d, dst->d, nwords));
  KPM_HIP_TRY(hipEventRecord(ctx->events[15], ctx->stream));
  KPM_HIP_TRY(hipEventSynchronize(ctx->events[15]));
  KPM_HIP_TRY(hipEventElapsedTime(ms, ctx->events[14], ctx->events[15]));
  return KPM_OK;
}

}  // extern "C"
