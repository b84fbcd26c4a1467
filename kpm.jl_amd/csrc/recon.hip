// recon.hip -- FP64 reconstruction heads on the device.
//
//   kpm_cheb_lin_trans     y(x) = sum_j mu~_j T_{n_j}(x)
//                          src/utils/chebyshev_lintrans.jl:12-27 (the device
//                          method of ext/KPMCUDAExt.jl:207-236), consumed by
//                          dos (src/applications/dos.jl:128-133)
//   kpm_kubo_bastin_nodes  the node contraction of kubo_bastin_cond
//                          (src/applications/conductivity.jl:133-155):
//     acc = sum_k phi_k [ sum_m U[m,k] WM[m,k] + sum_n V[n,k] WN[n,k] ],
//     U = mu~^T C, V = mu~ C, C[n,k] = T_n(x_k) = cos(n th_k),
//     WM[m,k] = (x_k - i m s_k) e^{ i m th_k},
//     WN[n,k] = (x_k + i n s_k) e^{-i n th_k},   th_k = acos x_k, s_k = sqrt(1 - x_k^2).
//
// Both are FP64 throughout.  Every sum has a fixed order (lane-strided
// partials reduced by a fixed shuffle tree; per-workgroup partials summed on
// the host in workgroup order), so results are run-to-run identical.
#include <algorithm>
#include <vector>

#include "kpm_internal.h"

namespace kpm {

// One wave per x: lanes stride over the n grid, then a fixed xor tree.
__global__ void __launch_bounds__(256)
    k_lin_trans(const double* __restrict__ x, int64_t Nx, const int64_t* __restrict__ n_grid,
                const double2* __restrict__ mu, int64_t Nn, double2* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= Nx) return;  // uniform per wave
  const double th = acos(x[i]);
  double sr = 0.0, si = 0.0;
  for (int64_t j = lane; j < Nn; j += 64) {
    const double t = cos((double)n_grid[j] * th);
    const double2 m = mu[j];
    sr = fma(t, m.x, sr);
    si = fma(t, m.y, si);
  }
  for (int o = 32; o > 0; o >>= 1) {
    sr += __shfl_xor(sr, o);
    si += __shfl_xor(si, o);
  }
  if (lane == 0) y[i] = make_double2(sr, si);
}

// Tile: 32 output indices p x 64 nodes k, contraction over q in chunks of 32.
// TRANS = true:  out[p,k] = sum_q mu~[q,p] C[q,k]  (U, weight WM with index p = m)
// TRANS = false: out[p,k] = sum_q mu~[p,q] C[q,k]  (V, weight WN with index p = n)
// Thread t owns p = t / 8 and nodes k = (t % 8) + 8 j, j < 8.
constexpr int kTP = 32, kTK = 64, kTQ = 32;

template <bool TRANS>
__global__ void __launch_bounds__(256)
    k_kubo_nodes(const double2* __restrict__ mu, int NC, const double* __restrict__ xk,
                 const double* __restrict__ phi, int K, double2* __restrict__ part) {
  __shared__ double2 As[kTP][kTQ + 1];
  __shared__ double Cs[kTQ][kTK];
  __shared__ double th_s[kTK];
  __shared__ double2 red[256];
  const int t = threadIdx.x;
  const int p0 = blockIdx.x * kTP;
  const int k0 = blockIdx.y * kTK;
  if (t < kTK) th_s[t] = k0 + t < K ? acos(xk[k0 + t]) : 0.0;
  const int pl = t >> 3, kl = t & 7;
  double ar[8], ai[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) ar[j] = ai[j] = 0.0;
  __syncthreads();
  for (int q0 = 0; q0 < NC; q0 += kTQ) {
    // stage A (32 p x 32 q complex) and C (32 q x 64 k real)
    for (int e = t; e < kTP * kTQ; e += 256) {
      const int pp = e / kTQ, qq = e % kTQ;
      const int p = p0 + pp, q = q0 + qq;
      double2 v = make_double2(0.0, 0.0);
      if (p < NC && q < NC) v = TRANS ? mu[(size_t)q + (size_t)NC * p] : mu[(size_t)p + (size_t)NC * q];
      As[pp][qq] = v;
    }
    for (int e = t; e < kTQ * kTK; e += 256) {
      const int qq = e / kTK, kk = e % kTK;
      Cs[qq][kk] = cos((double)(q0 + qq) * th_s[kk]);
    }
    __syncthreads();
#pragma unroll 4
    for (int qq = 0; qq < kTQ; ++qq) {
      const double2 a = As[pl][qq];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const double c = Cs[qq][kl + 8 * j];
        ar[j] = fma(a.x, c, ar[j]);
        ai[j] = fma(a.y, c, ai[j]);
      }
    }
    __syncthreads();
  }
  // epilogue: weight, node factor, fixed-order reduction over the tile
  double sr = 0.0, si = 0.0;
  const int p = p0 + pl;
  if (p < NC) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int kk = kl + 8 * j;
      const int k = k0 + kk;
      if (k < K) {
        const double th = th_s[kk];
        const double x = xk[k];
        const double s = sqrt(1.0 - x * x);
        double sn, cs;
        sincos((double)p * th, &sn, &cs);
        // TRANS (WM): (x - i p s)(cs + i sn);  !TRANS (WN): (x + i p s)(cs - i sn)
        const double ps = TRANS ? -(double)p * s : (double)p * s;
        const double sg = TRANS ? sn : -sn;
        const double wr = x * cs - ps * sg;
        const double wi = x * sg + ps * cs;
        const double f = phi[k];
        sr = fma(f, ar[j] * wr - ai[j] * wi, sr);
        si = fma(f, ar[j] * wi + ai[j] * wr, si);
      }
    }
  }
  red[t] = make_double2(sr, si);
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      red[t].x += red[t + w].x;
      red[t].y += red[t + w].y;
    }
    __syncthreads();
  }
  if (t == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_cheb_lin_trans(kpm_ctx* ctx, int64_t Nx, const double* x, int64_t Nn, const int64_t* n_grid,
                       const double* mu_tilde, double* y) {
  if (!ctx || (Nx > 0 && (!x || !y)) || (Nn > 0 && (!n_grid || !mu_tilde)))
    return fail(KPM_EINVAL, "null argument");
  if (Nx < 0 || Nn < 0) return fail(KPM_EINVAL, "negative grid size");
  if (Nx == 0) return KPM_OK;
  for (int64_t i = 0; i < Nx; ++i)
    if (!(x[i] >= -1.0 && x[i] <= 1.0)) return fail(KPM_EINVAL, "x must lie in [-1, 1]");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const size_t bx = (size_t)Nx * 8, bn = (size_t)std::max<int64_t>(Nn, 1) * 8,
               bm = (size_t)std::max<int64_t>(Nn, 1) * 16, by = (size_t)Nx * 16;
  KPM_TRY(ensure(ctx, ctx->staging, bx + bn + bm + by));
  char* base = (char*)ctx->staging.ptr;
  double* dx = (double*)base;
  int64_t* dn = (int64_t*)(base + bx);
  double2* dm = (double2*)(base + bx + bn);
  double2* dy = (double2*)(base + bx + bn + bm);
  KPM_HIP_TRY(hipMemcpyAsync(dx, x, bx, hipMemcpyHostToDevice, ctx->stream));
  if (Nn > 0) {
    KPM_HIP_TRY(hipMemcpyAsync(dn, n_grid, (size_t)Nn * 8, hipMemcpyHostToDevice, ctx->stream));
    KPM_HIP_TRY(hipMemcpyAsync(dm, mu_tilde, (size_t)Nn * 16, hipMemcpyHostToDevice, ctx->stream));
  }
  hipLaunchKernelGGL(k_lin_trans, dim3((unsigned)((Nx + 3) / 4)), dim3(256), 0, ctx->stream, dx, Nx,
                     dn, dm, Nn, dy);
  KPM_HIP_TRY(hipGetLastError());
  KPM_HIP_TRY(hipMemcpyAsync(y, dy, by, hipMemcpyDeviceToHost, ctx->stream));
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return KPM_OK;
}

int kpm_kubo_bastin_nodes(kpm_ctx* ctx, int64_t NC, const double* mu_tilde, int64_t K,
                          const double* x, const double* phi, double* acc) {
  if (!ctx || !mu_tilde || !acc || (K > 0 && (!x || !phi))) return fail(KPM_EINVAL, "null argument");
  if (NC < 1 || K < 0) return fail(KPM_EINVAL, "NC must be >= 1 and K >= 0");
  if (NC > (1 << 20) || K > INT32_MAX / 2) return fail(KPM_EINVAL, "problem too large");
  acc[0] = acc[1] = 0.0;
  if (K == 0) return KPM_OK;
  for (int64_t k = 0; k < K; ++k)
    if (!(x[k] > -1.0 && x[k] < 1.0)) return fail(KPM_EINVAL, "nodes must lie in (-1, 1)");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int64_t gp = (NC + kTP - 1) / kTP, gk = (K + kTK - 1) / kTK;
  if (gk > 65535) return fail(KPM_EINVAL, "too many quadrature nodes (at most 4,194,240)");
  const size_t bmu = (size_t)(NC * NC) * 16, bx = (size_t)K * 8, bp = (size_t)(2 * gp * gk) * 16;
  KPM_TRY(ensure(ctx, ctx->staging, bmu + 2 * bx + bp));
  char* base = (char*)ctx->staging.ptr;
  double2* dmu = (double2*)base;
  double* dx = (double*)(base + bmu);
  double* dphi = (double*)(base + bmu + bx);
  double2* dpart = (double2*)(base + bmu + 2 * bx);
  KPM_HIP_TRY(hipMemcpyAsync(dmu, mu_tilde, bmu, hipMemcpyHostToDevice, ctx->stream));
  KPM_HIP_TRY(hipMemcpyAsync(dx, x, bx, hipMemcpyHostToDevice, ctx->stream));
  KPM_HIP_TRY(hipMemcpyAsync(dphi, phi, bx, hipMemcpyHostToDevice, ctx->stream));
  dim3 grid((unsigned)gp, (unsigned)gk);
  hipLaunchKernelGGL(k_kubo_nodes<true>, grid, dim3(256), 0, ctx->stream, dmu, (int)NC, dx, dphi,
                     (int)K, dpart);
  hipLaunchKernelGGL(k_kubo_nodes<false>, grid, dim3(256), 0, ctx->stream, dmu, (int)NC, dx, dphi,
                     (int)K, dpart + gp * gk);
  KPM_HIP_TRY(hipGetLastError());
  std::vector<double> h((size_t)(4 * gp * gk));
  KPM_HIP_TRY(hipMemcpyAsync(h.data(), dpart, bp, hipMemcpyDeviceToHost, ctx->stream));
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  double sr = 0.0, si = 0.0;
  for (size_t i = 0; i < h.size() / 2; ++i) {
    sr += h[2 * i];
    si += h[2 * i + 1];
  }
  acc[0] = sr;
  acc[1] = si;
  return KPM_OK;
}

}  // extern "C"
