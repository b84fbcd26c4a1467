// lanczos.hip -- FP64 spectral-edge estimate of a Hermitian operator on the
// device, for normalizeH (src/utils/external.jl:41-76).  The reference calls
// Arpack eigs (:LR / :SR / :LM, tol 1e-3, maxiter 300) on the host; here the
// Lanczos three-term recurrence runs on the device with the fused pass:
//
//   w = H v_j - beta_{j-1} v_{j-1}        (k_step, MODE_DBL: alpha_j = <v_j|w>)
//   w -= alpha_j v_j,  beta_j = |w|,  v_{j+1} = w / beta_j
//
// The extreme eigenvalues of the tridiagonal T_j (Sturm-sequence bisection
// on the host, O(j) per probe) are checked every 10 steps; the run stops
// when both edges move by less than tol * max(|emin|, |emax|) between checks,
// when beta_j vanishes (an invariant subspace: the Ritz values are exact), or
// after max_iter steps.  No reorthogonalisation: lost orthogonality only
// duplicates converged Ritz values, it never moves the extreme ones outward.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "kpm_internal.h"

namespace kpm {

__global__ void k_axpy_real(double2* __restrict__ w, const double2* __restrict__ v, double a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 x = v[i];
    double2 y = w[i];
    y.x = fma(-a, x.x, y.x);
    y.y = fma(-a, x.y, y.y);
    w[i] = y;
  }
}

__global__ void k_scale_copy(double2* __restrict__ dst, const double2* __restrict__ src, double s, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 x = src[i];
    dst[i] = make_double2(x.x * s, x.y * s);
  }
}

// Number of eigenvalues of the symmetric tridiagonal (a, b) below x.
static int64_t sturm_count(const std::vector<double>& a, const std::vector<double>& b, double x) {
  int64_t cnt = 0;
  double q = 1.0;
  for (size_t i = 0; i < a.size(); ++i) {
    const double off = i == 0 ? 0.0 : b[i - 1] * b[i - 1];
    q = (a[i] - x) - (i == 0 ? 0.0 : off / q);
    if (q == 0.0) q = -1e-300;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

// Smallest and largest eigenvalue of the tridiagonal by bisection inside the
// Gershgorin interval.
static void tridiag_edges(const std::vector<double>& a, const std::vector<double>& b, double* lo,
                          double* hi) {
  const size_t m = a.size();
  double glo = INFINITY, ghi = -INFINITY;
  for (size_t i = 0; i < m; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
    glo = std::min(glo, a[i] - r);
    ghi = std::max(ghi, a[i] + r);
  }
  const double span = std::max(ghi - glo, 1e-300);
  auto bisect = [&](int64_t k) {  // k-th smallest eigenvalue (0-based)
    double l = glo - 1e-12 * span, h = ghi + 1e-12 * span;
    for (int it = 0; it < 200 && h - l > 1e-15 * std::max(1.0, std::fabs(l) + std::fabs(h)); ++it) {
      const double mid = 0.5 * (l + h);
      if (sturm_count(a, b, mid) > k)
        h = mid;
      else
        l = mid;
    }
    return 0.5 * (l + h);
  };
  *lo = bisect(0);
  *hi = bisect((int64_t)m - 1);
}

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_spectral_edges(kpm_ctx* ctx, const kpm_csr* H, int64_t max_iter, double tol, uint64_t seed,
                       double* emin, double* emax, int64_t* iters) {
  if (!ctx || !H || !emin || !emax) return fail(KPM_EINVAL, "null argument");
  if (max_iter < 1) return fail(KPM_EINVAL, "max_iter must be >= 1");
  if (!(tol > 0.0 && tol < 1.0)) return fail(KPM_EINVAL, "tol must be in (0, 1)");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int64_t NH = H->n;
  // start vector: splitmix64-driven uniform values in [-0.5, 0.5), normalised
  std::vector<double> v0((size_t)(2 * NH), 0.0);
  uint64_t s = seed;
  double nrm = 0.0;
  for (int64_t i = 0; i < NH; ++i) {
    s += 0x9E3779B97F4A7C15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    v0[2 * i] = u;
    nrm += u * u;
  }
  nrm = std::sqrt(nrm);
  for (int64_t i = 0; i < NH; ++i) v0[2 * i] /= nrm;

  kpm_block blk[3];
  for (int q = 0; q < 3; ++q) KPM_TRY(work_block(ctx, ctx->work[q], NH, 1, KPM_VEC_F64, &blk[q]));
  const size_t hbytes = (size_t)NH * 16;
  KPM_TRY(ensure(ctx, ctx->staging, hbytes));
  KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, v0.data(), hbytes, hipMemcpyHostToDevice, ctx->stream));
  KPM_TRY(launch_upload(ctx, ctx->staging.ptr, &blk[0]));
  KPM_TRY(ensure(ctx, ctx->mu, 2 * sizeof(double2)));
  const int64_t nwords = NH * blk[0].NRp;
  const unsigned g = (unsigned)std::min<int64_t>((nwords + 255) / 256, 4096);

  std::vector<double> al, be;  // T_j diagonal and off-diagonal
  int vp = 2, vc = 0, vw = 1;  // v_{j-1}, v_j, w
  double beta_prev = 0.0, lo = 0.0, hi = 0.0, plo = NAN, phi = NAN;
  int64_t j = 0;
  for (; j < max_iter; ++j) {
    int G = 0;
    StepLaunch st{H, &blk[vc], j > 0 ? &blk[vp] : nullptr, &blk[vw], nullptr, 1.0, -beta_prev, MODE_DBL};
    if (j == 0) st.beta = 0.0;
    KPM_TRY(launch_step(ctx, st, StepConfig{}, &G));
    Combine c{2, {0, 1}, {1.0, 1.0}, {-1, -1}, {0.0, 0.0}};
    KPM_TRY(launch_combine(ctx, G, 1, blk[vc].NRp, c, ctx->mu.ptr, 2));
    double d[4];
    KPM_HIP_TRY(hipMemcpyAsync(d, ctx->mu.ptr, sizeof d, hipMemcpyDeviceToHost, ctx->stream));
    KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const double alpha = d[2];  // Re <v_j | w>
    al.push_back(alpha);
    hipLaunchKernelGGL(k_axpy_real, dim3(g), dim3(256), 0, ctx->stream, (double2*)blk[vw].d,
                       (const double2*)blk[vc].d, alpha, nwords);
    KPM_HIP_TRY(hipGetLastError());
    KPM_TRY(launch_dot(ctx, &blk[vw], &blk[vw], &G));
    Combine cn{1, {0, 0}, {1.0, 0.0}, {-1, -1}, {0.0, 0.0}};
    KPM_TRY(launch_combine(ctx, G, 1, blk[vw].NRp, cn, ctx->mu.ptr, 1));
    KPM_HIP_TRY(hipMemcpyAsync(d, ctx->mu.ptr, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const double beta = std::sqrt(std::max(d[0], 0.0));
    const bool last = j + 1 == max_iter || j + 1 == NH;
    const bool exhausted = !(beta > 1e-14 * std::max(1.0, std::fabs(alpha)));
    if (last || exhausted || (j + 1) % 10 == 0) {
      tridiag_edges(al, be, &lo, &hi);
      const double scale = std::max(std::fabs(lo), std::fabs(hi));
      const bool settled = j + 1 >= 20 && std::fabs(lo - plo) <= tol * scale &&
                           std::fabs(hi - phi) <= tol * scale;
      plo = lo;
      phi = hi;
      if (last || exhausted || settled) {
        ++j;
        break;
      }
    }
    be.push_back(beta);
    // v_{j+1} = w / beta into the v_{j-1} buffer; rotate
    hipLaunchKernelGGL(k_scale_copy, dim3(g), dim3(256), 0, ctx->stream, (double2*)blk[vp].d,
                       (const double2*)blk[vw].d, 1.0 / beta, nwords);
    KPM_HIP_TRY(hipGetLastError());
    const int old_prev = vp;
    vp = vc;
    vc = old_prev;
    beta_prev = beta;
  }
  *emin = lo;
  *emax = hi;
  if (iters) *iters = j;
  return KPM_OK;
}

}  // extern "C"
