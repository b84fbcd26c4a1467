// adjoint.hip -- reverse-mode gradient of the equal-vector moments.
//
// No counterpart in the reference (its moments are not differentiable through
// the recurrence).  For mu_n = (1/NR) sum_r <v_r|T_n(H)|v_r> (n < NC) and
// g_n = dL/dmu_n, the derivative of L with respect to the stored entries of
// H is, on H's sparsity pattern,
//
//   M_ab = (1/NR) sum_r sum_{j=1}^{NC-1} c_j conj(beta_j[a,r]) alpha_{j-1}[b,r],
//   c_1 = 1, c_j = 2 (j >= 2),   so that dL = sum_ab dH_ab M_ab,
//
// with alpha_k = T_k(H) v the forward Chebyshev vectors and beta_j the reverse
// Clenshaw vectors beta_j = g_j v + 2 H beta_{j+1} - beta_{j+2},
// beta_NC = beta_{NC+1} = 0.  (dalpha obeys the recurrence driven by the
// sources dH alpha_0 and 2 dH alpha_{j-1}; its Green's function is U_{n-j}(H),
// and beta_j = sum_{n>=j} g_n U_{n-j}(H) v collects it backwards.)
//
// alpha_{j-1} is needed in reverse order: the forward pass keeps the pairs
// (alpha_{sK}, alpha_{sK+1}) every K ~ sqrt(NC) steps, and the reverse pass
// recomputes one segment of K vectors at a time from its pair.
// Memory: (2 ceil(NC/K) + K + 5) blocks; work: ~3x a forward recurrence.
#include <algorithm>
#include <cmath>
#include <vector>

#include "step_kernel.h"

namespace kpm {

// M[k] += c * sum_{r} conj(beta[a,r]) alpha[colind[k], r] for every stored
// entry k of row a; LPR lanes per row, reduced across the row's lanes.
template <int PREC, int LPR>
__global__ void __launch_bounds__(kBlock)
    k_sddmm(const int* __restrict__ rowptr, const int* __restrict__ colind,
            const typename Word<PREC>::T* __restrict__ beta, const typename Word<PREC>::T* __restrict__ alpha,
            double2* __restrict__ M, int NH, int RW, double c) {
  using WT = Word<PREC>;
  constexpr int CPL = WT::CPL;
  constexpr int RPW = 64 / LPR;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int sub = lane % LPR;
  const int rslot = lane / LPR;
  const int wcol = blockIdx.y * LPR + sub;
  const int rows_per_block = 4 * RPW;
  for (int base = blockIdx.x * rows_per_block; base < NH; base += gridDim.x * rows_per_block) {
    const int row = base + wave * RPW + rslot;
    const bool live = row < NH;
    const int rc = live ? row : 0;
    double br[CPL], bi[CPL];
    WT::unpack(beta[(size_t)rc * RW + wcol], br, bi);
    const int s0 = rowptr[rc];
    const int n = live ? rowptr[rc + 1] - s0 : 0;
    // the row's lanes agree on n; the wave loops to its longest row
    int nmax = n;
    for (int o = 1; o < 64; o <<= 1) nmax = max(nmax, __shfl_xor(nmax, o));
    for (int j = 0; j < nmax; ++j) {
      const bool v = j < n;
      const int k = v ? s0 + j : 0;
      const int col = v ? colind[k] : 0;
      double ar[CPL], ai[CPL];
      WT::unpack(alpha[(size_t)col * RW + wcol], ar, ai);
      double pr = 0.0, pi = 0.0;
#pragma unroll
      for (int q = 0; q < CPL; ++q) {
        pr = fma(br[q], ar[q], fma(bi[q], ai[q], pr));
        pi = fma(br[q], ai[q], fma(-bi[q], ar[q], pi));
      }
#pragma unroll
      for (int o = 1; o < LPR; o <<= 1) {
        pr += __shfl_xor(pr, o);
        pi += __shfl_xor(pi, o);
      }
      if (v && sub == 0 && blockIdx.y == 0 && gridDim.y == 1) {
        double2 m = M[k];
        m.x = fma(c, pr, m.x);
        m.y = fma(c, pi, m.y);
        M[k] = m;
      }
      if (gridDim.y > 1 && v && sub == 0) {
        // rows wider than 64 words: each row chunk adds its own columns
        atomicAdd(&M[k].x, c * pr);
        atomicAdd(&M[k].y, c * pi);
      }
    }
  }
}

template <int PREC>
static int launch_sddmm_t(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* beta,
                          const kpm_block* alpha, double2* M, double c) {
  const int RW = row_words(beta->NR, PREC);
  const int lpr = std::min(RW, 64);
  const int chunks = RW / lpr;
  const int64_t rows_per_block = 4LL * (64 / lpr);
  const int64_t nb = (beta->NH + rows_per_block - 1) / rows_per_block;
  const unsigned G = (unsigned)std::min<int64_t>(nb, (int64_t)std::max(ctx->num_cus, 8) * 8);
  using T = typename Word<PREC>::T;
#define KPM_SDDMM(L)                                                                              \
  case L:                                                                                         \
    hipLaunchKernelGGL((k_sddmm<PREC, L>), dim3(G, chunks), dim3(kBlock), 0, ctx->stream,         \
                       A->rowptr, A->colind, (const T*)beta->d, (const T*)alpha->d, M,             \
                       (int)beta->NH, RW, c);                                                     \
    break;
  switch (lpr) {
    KPM_SDDMM(1)
    KPM_SDDMM(2)
    KPM_SDDMM(4)
    KPM_SDDMM(8)
    KPM_SDDMM(16)
    KPM_SDDMM(32)
    default:
      hipLaunchKernelGGL((k_sddmm<PREC, 64>), dim3(G, chunks), dim3(kBlock), 0, ctx->stream,
                         A->rowptr, A->colind, (const T*)beta->d, (const T*)alpha->d, M,
                         (int)beta->NH, RW, c);
  }
#undef KPM_SDDMM
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

static int launch_sddmm(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* beta, const kpm_block* alpha,
                        double2* M, double c) {
  return beta->prec == KPM_VEC_F32 ? launch_sddmm_t<KPM_VEC_F32>(ctx, A, beta, alpha, M, c)
                                   : launch_sddmm_t<KPM_VEC_F64>(ctx, A, beta, alpha, M, c);
}

static int run_step(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* x, const kpm_block* z,
                    kpm_block* out, const kpm_block* src, double alpha, double beta, double gamma,
                    int mode) {
  StepLaunch s{A, x, z, out, src, alpha, beta, mode, gamma};
  return launch_step(ctx, s, tuned_config(ctx, A, x->NR, x->prec, MODE_NONE), nullptr);
}

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_1d_adjoint(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t NR, int64_t NH,
                   const double* psi, const double* g, double* grad, int vec_prec,
                   int64_t ckpt_stride) {
  if (!ctx || !H || !psi || !g || !grad) return fail(KPM_EINVAL, "null argument");
  if (NC < 2 || NR <= 0 || NH != H->n) return fail(KPM_EINVAL, "bad sizes");
  if (vec_prec != KPM_VEC_F64 && vec_prec != KPM_VEC_F32) return fail(KPM_EINVAL, "bad vec_prec");
  if ((int64_t)H->csc_to_csr.size() != H->nnz)
    return fail(KPM_EINVAL, "operator was not built from CSC arrays (kpm_csr_from_csc)");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int64_t K = ckpt_stride > 0 ? ckpt_stride : (int64_t)std::ceil(std::sqrt((double)NC));
  const int64_t nalpha = NC - 1;               // alpha_0 .. alpha_{NC-2} are used
  const int64_t nseg = (nalpha + K - 1) / K;   // segments of K alphas
  const int64_t NRp = padded_nr(NR, vec_prec);
  const size_t bbytes = (size_t)(NH * NRp) * elem_bytes(vec_prec);
  // workspace: v, 2 forward ring, 2 beta ring, 2*nseg checkpoints, K segment
  const int64_t nblk = 5 + 2 * nseg + K;
  Buffer ws;
  KPM_TRY(ensure(ctx, ws, bbytes * (size_t)nblk));
  auto blk = [&](int64_t i) {
    kpm_block b;
    b.ctx = ctx;
    b.NH = NH;
    b.NR = NR;
    b.NRp = NRp;
    b.prec = vec_prec;
    b.d = (char*)ws.ptr + bbytes * (size_t)i;
    b.owned = false;
    return b;
  };
  kpm_block v = blk(0), f0 = blk(1), f1 = blk(2), b0 = blk(3), b1 = blk(4);
  auto ckA = [&](int64_t s) { return blk(5 + 2 * s); };
  auto ckB = [&](int64_t s) { return blk(6 + 2 * s); };
  auto seg = [&](int64_t i) { return blk(5 + 2 * nseg + i); };
  Buffer Mbuf;
  int rc = KPM_OK;
  do {
    if ((rc = ensure(ctx, Mbuf, (size_t)std::max<int64_t>(H->nnz, 1) * sizeof(double2))) != KPM_OK) break;
    KPM_HIP_TRY(hipMemsetAsync(Mbuf.ptr, 0, (size_t)std::max<int64_t>(H->nnz, 1) * sizeof(double2), ctx->stream));
    KPM_HIP_TRY(hipMemsetAsync(ws.ptr, 0, bbytes * (size_t)nblk, ctx->stream));
    if ((rc = ensure(ctx, ctx->staging, (size_t)(NH * NR) * 16)) != KPM_OK) break;
    KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, psi, (size_t)(NH * NR) * 16, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = launch_upload(ctx, ctx->staging.ptr, &v)) != KPM_OK) break;
    // forward: alpha_0 = v, alpha_1 = H v, checkpoint pairs (alpha_{sK}, alpha_{sK+1})
    kpm_block ring[2] = {f0, f1};
    if ((rc = kpm_block_copy(&ring[0], &v)) != KPM_OK) break;
    if ((rc = run_step(ctx, H, &ring[0], nullptr, &ring[1], nullptr, 1.0, 0.0, 0.0, MODE_NONE)) != KPM_OK) break;
    const int64_t klast = (nseg - 1) * K + 1;
    for (int64_t k = 1; k <= klast && rc == KPM_OK; ++k) {
      if (k >= 2)
        rc = run_step(ctx, H, &ring[(k - 1) % 2], &ring[k % 2], &ring[k % 2], nullptr, 2.0, -1.0, 0.0, MODE_NONE);
      if (rc == KPM_OK && (k - 1) % K == 0) {  // pair (alpha_{k-1}, alpha_k) starts segment (k-1)/K
        const int64_t s = (k - 1) / K;
        kpm_block a = ckA(s), b = ckB(s);
        rc = kpm_block_copy(&a, &ring[(k - 1) % 2]);
        if (rc == KPM_OK) rc = kpm_block_copy(&b, &ring[k % 2]);
      }
    }
    if (rc != KPM_OK) break;
    // reverse: beta ring starts at beta_NC = beta_{NC+1} = 0 (b0, b1 zeroed)
    kpm_block bring[2] = {b0, b1};  // bring[j % 2] holds beta_j
    for (int64_t s = nseg - 1; s >= 0 && rc == KPM_OK; --s) {
      const int64_t lo = s * K, hi = std::min(lo + K - 1, nalpha - 1);
      kpm_block s0 = seg(0), a = ckA(s), b = ckB(s);
      rc = kpm_block_copy(&s0, &a);
      if (rc == KPM_OK && hi > lo) {
        kpm_block s1 = seg(1);
        rc = kpm_block_copy(&s1, &b);
      }
      for (int64_t i = lo + 2; i <= hi && rc == KPM_OK; ++i) {
        kpm_block si = seg(i - lo), sm1 = seg(i - lo - 1), sm2 = seg(i - lo - 2);
        rc = run_step(ctx, H, &sm1, &sm2, &si, nullptr, 2.0, -1.0, 0.0, MODE_NONE);
      }
      for (int64_t i = hi; i >= lo && rc == KPM_OK; --i) {
        const int64_t j = i + 1;  // beta_j pairs with alpha_{j-1} = alpha_i
        // beta_j = g_j v + 2 H beta_{j+1} - beta_{j+2}, written over beta_{j+2}
        kpm_block& bj = bring[j % 2];
        rc = run_step(ctx, H, &bring[(j + 1) % 2], &bj, &bj, &v, 2.0, -1.0, g[j], MODE_SRC);
        if (rc == KPM_OK) {
          kpm_block ai = seg(i - lo);
          rc = launch_sddmm(ctx, H, &bj, &ai, (double2*)Mbuf.ptr, j == 1 ? 1.0 : 2.0);
        }
      }
    }
    if (rc != KPM_OK) break;
    std::vector<double> m(2 * std::max<int64_t>(H->nnz, 1));
    KPM_HIP_TRY(hipMemcpyAsync(m.data(), Mbuf.ptr, (size_t)std::max<int64_t>(H->nnz, 1) * sizeof(double2),
                               hipMemcpyDeviceToHost, ctx->stream));
    KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const double inv = 1.0 / (double)NR;
    for (int64_t k = 0; k < H->nnz; ++k) {
      const int64_t p = H->csc_to_csr[k];
      grad[2 * k] = m[2 * p] * inv;
      grad[2 * k + 1] = m[2 * p + 1] * inv;
    }
  } while (false);
  (void)hipStreamSynchronize(ctx->stream);
  if (Mbuf.ptr) (void)hipFree(Mbuf.ptr);
  if (ws.ptr) (void)hipFree(ws.ptr);
  return rc;
}

}  // extern "C"
