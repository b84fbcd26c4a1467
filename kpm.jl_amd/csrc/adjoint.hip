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
// alpha_{j-1} is needed in reverse order.  When the NC-1 forward vectors fit
// in HBM (ckpt_stride 0 or 1; 69 GB for NR = 16 at NH = 262,144, NC = 1024)
// they are all kept and nothing is recomputed.  Otherwise the forward pass
// keeps the pairs (alpha_{sK}, alpha_{sK+1}) every K ~ sqrt(NC) steps and
// the reverse pass recomputes one segment of K vectors at a time from its
// pair: (2 ceil(NC/K) + K + 5) blocks and ~one more forward recurrence.
//
// The gradient accumulation (an SDDMM on H's pattern) runs once per window of
// up to kWin reverse steps: the last kWin beta vectors stay in a ring and one
// pass adds all their terms, so the nnz-sized accumulator is read and written
// once per window instead of once per step.
#include <algorithm>
#include <cmath>
#include <vector>

#include "step_kernel.h"

namespace kpm {


template <int PREC>
struct SddmmWindow {
  const typename Word<PREC>::T* beta[kWin];
  const typename Word<PREC>::T* alpha[kWin];
  double c[kWin];
  int W;
};

// M[k] += sum_{w < W} c_w sum_r conj(beta_w[a,r]) alpha_w[colind[k], r] for
// every stored entry k of row a; LPR lanes per row, reduced across the row's
// lanes (fixed xor order), then across row chunks with atomics only when a
// row is wider than 64 words.
template <int PREC, int LPR>
__global__ void __launch_bounds__(kBlock)
    k_sddmm(const int* __restrict__ rowptr, const int* __restrict__ colind, SddmmWindow<PREC> win,
            double2* __restrict__ M, int NH, int RW) {
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
    double br[kWin][CPL], bi[kWin][CPL];
#pragma unroll
    for (int w = 0; w < kWin; ++w) {
      if (w < win.W) {
        WT::unpack(win.beta[w][(size_t)rc * RW + wcol], br[w], bi[w]);
      } else {
#pragma unroll
        for (int q = 0; q < CPL; ++q) br[w][q] = bi[w][q] = 0.0;
      }
    }
    const int s0 = rowptr[rc];
    const int n = live ? rowptr[rc + 1] - s0 : 0;
    // the row's lanes agree on n; the wave loops to its longest row
    int nmax = n;
    for (int o = 1; o < 64; o <<= 1) nmax = max(nmax, __shfl_xor(nmax, o));
    for (int j = 0; j < nmax; ++j) {
      const bool v = j < n;
      const int k = v ? s0 + j : 0;
      const int col = v ? colind[k] : 0;
      double pr = 0.0, pi = 0.0;
#pragma unroll
      for (int w = 0; w < kWin; ++w) {
        if (w < win.W) {
          double ar[CPL], ai[CPL];
          WT::unpack(win.alpha[w][(size_t)col * RW + wcol], ar, ai);
          double qr = 0.0, qi = 0.0;
#pragma unroll
          for (int q = 0; q < CPL; ++q) {
            qr = fma(br[w][q], ar[q], fma(bi[w][q], ai[q], qr));
            qi = fma(br[w][q], ai[q], fma(-bi[w][q], ar[q], qi));
          }
          pr = fma(win.c[w], qr, pr);
          pi = fma(win.c[w], qi, pi);
        }
      }
#pragma unroll
      for (int o = 1; o < LPR; o <<= 1) {
        pr += __shfl_xor(pr, o);
        pi += __shfl_xor(pi, o);
      }
      if (v && sub == 0 && gridDim.y == 1) {
        double2 m = M[k];
        m.x += pr;
        m.y += pi;
        M[k] = m;
      }
      if (gridDim.y > 1 && v && sub == 0) {
        // rows wider than 64 words: each row chunk adds its own columns
        atomicAdd(&M[k].x, pr);
        atomicAdd(&M[k].y, pi);
      }
    }
  }
}

// The same sums for rows of LPR >= 2 lanes, at most KB <= 16 nonzeros and a
// single row chunk, with every memory access of a row issued together: with
// G = min(LPR, 16), lane j of a row holds column index and accumulator entry
// of nonzeros j, j + G, ... (one coalesced load each), group_bcast hands slot
// jj's column to the row, the window's KB gathers per step go out at once,
// and after the row reduction (the xor order above) each lane adds its own
// entries and stores them coalesced.  Per entry the terms are summed in the
// order of k_sddmm, so both kernels give bitwise the same M.
template <int PREC, int LPR, int KB>
__global__ void __launch_bounds__(kBlock)
    k_sddmm_rows(const int* __restrict__ rowptr, const int* __restrict__ colind,
                 SddmmWindow<PREC> win, double2* __restrict__ M, int NH, int RW) {
  using WT = Word<PREC>;
  constexpr int CPL = WT::CPL;
  constexpr int RPW = 64 / LPR;
  constexpr int MG = LPR >= 16 ? 16 : LPR;
  constexpr int ME = (KB + MG - 1) / MG;
  static_assert(MG >= 2 && KB <= 16, "rows of 2+ lanes, at most 16 nonzeros");
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int sub = lane % LPR;
  const int rslot = lane / LPR;
  const int j = lane & (MG - 1);
  const bool owner = sub < MG;  // rows wider than 16 lanes hold each entry once per DPP row
  const int rows_per_block = 4 * RPW;
  for (int base = blockIdx.x * rows_per_block; base < NH; base += gridDim.x * rows_per_block) {
    const int row = base + wave * RPW + rslot;
    const bool live = row < NH;
    const int rc = live ? row : 0;
    const int s0 = rowptr[rc];
    const int n = live ? rowptr[rc + 1] - s0 : 0;
    int cl[ME];
    double2 m[ME];
#pragma unroll
    for (int e = 0; e < ME; ++e) {
      const int idx = e * MG + j;
      const bool v = idx < n;
      const int k = v ? s0 + idx : 0;
      const int c = colind[k];
      const double2 mv = M[k];
      cl[e] = v ? c : 0;
      m[e] = mv;
    }
    double br[kWin][CPL], bi[kWin][CPL];
#pragma unroll
    for (int w = 0; w < kWin; ++w) {
      if (w < win.W) {
        WT::unpack(win.beta[w][(size_t)rc * RW + sub], br[w], bi[w]);
      } else {
#pragma unroll
        for (int q = 0; q < CPL; ++q) br[w][q] = bi[w][q] = 0.0;
      }
    }
    int col[KB];
    static_for<0, KB>([&](auto J) {
      constexpr int jj = decltype(J)::value;
      col[jj] = group_bcast<MG, jj % MG>(cl[jj / MG]);
    });
    double pr[KB], pi[KB];
#pragma unroll
    for (int jj = 0; jj < KB; ++jj) pr[jj] = pi[jj] = 0.0;
#pragma unroll
    for (int w = 0; w < kWin; ++w) {
      if (w < win.W) {  // uniform
        typename WT::T av[KB];
#pragma unroll
        for (int jj = 0; jj < KB; ++jj) av[jj] = win.alpha[w][(size_t)col[jj] * RW + sub];
#pragma unroll
        for (int jj = 0; jj < KB; ++jj) {
          double ar[CPL], ai[CPL];
          WT::unpack(av[jj], ar, ai);
          double qr = 0.0, qi = 0.0;
#pragma unroll
          for (int q = 0; q < CPL; ++q) {
            qr = fma(br[w][q], ar[q], fma(bi[w][q], ai[q], qr));
            qi = fma(br[w][q], ai[q], fma(-bi[w][q], ar[q], qi));
          }
          pr[jj] = fma(win.c[w], qr, pr[jj]);
          pi[jj] = fma(win.c[w], qi, pi[jj]);
        }
      }
    }
    static_for<0, KB>([&](auto J) {
      constexpr int jj = decltype(J)::value;
      double tr = pr[jj], ti = pi[jj];
#pragma unroll
      for (int o = 1; o < LPR; o <<= 1) {
        tr += __shfl_xor(tr, o);
        ti += __shfl_xor(ti, o);
      }
      if (j == jj % MG) {
        m[jj / MG].x += tr;
        m[jj / MG].y += ti;
      }
    });
#pragma unroll
    for (int e = 0; e < ME; ++e) {
      const int idx = e * MG + j;
      if (owner && idx < n) M[s0 + idx] = m[e];
    }
  }
}

template <int PREC>
static int launch_sddmm_t(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* const* beta,
                          const kpm_block* const* alpha, const double* c, int W, double2* M) {
  using T = typename Word<PREC>::T;
  SddmmWindow<PREC> win{};
  for (int w = 0; w < W; ++w) {
    win.beta[w] = (const T*)beta[w]->d;
    win.alpha[w] = (const T*)alpha[w]->d;
    win.c[w] = c[w];
  }
  win.W = W;
  const int RW = row_words(beta[0]->NR, PREC);
  const int lpr = std::min(RW, 64);
  const int chunks = RW / lpr;
  const int64_t rows_per_block = 4LL * (64 / lpr);
  const int64_t nb = (beta[0]->NH + rows_per_block - 1) / rows_per_block;
  const unsigned G = (unsigned)std::min<int64_t>(nb, (int64_t)std::max(ctx->num_cus, 8) * 8);
  const int w = A->max_row_nnz;
  if (lpr >= 2 && chunks == 1 && w <= 16) {
    const int kb = w <= 4 ? 4 : (w <= 8 ? 8 : (w <= 12 ? 12 : 16));
#define KPM_SDDMM_ROWS(L, KBv)                                                                    \
  if (lpr == L && kb == KBv) {                                                                    \
    hipLaunchKernelGGL((k_sddmm_rows<PREC, L, KBv>), dim3(G), dim3(kBlock), 0, ctx->stream,       \
                       A->rowptr, A->colind, win, M, (int)beta[0]->NH, RW);                       \
    KPM_HIP_TRY(hipGetLastError());                                                               \
    return KPM_OK;                                                                                \
  }
#define KPM_SDDMM_KB(L) KPM_SDDMM_ROWS(L, 4) KPM_SDDMM_ROWS(L, 8) KPM_SDDMM_ROWS(L, 12) KPM_SDDMM_ROWS(L, 16)
    KPM_SDDMM_KB(2)
    KPM_SDDMM_KB(4)
    KPM_SDDMM_KB(8)
    KPM_SDDMM_KB(16)
    KPM_SDDMM_KB(32)
    KPM_SDDMM_KB(64)
#undef KPM_SDDMM_KB
#undef KPM_SDDMM_ROWS
    return fail(KPM_EINVAL, "sddmm: no row kernel for this shape");
  }
#define KPM_SDDMM(L)                                                                              \
  case L:                                                                                         \
    hipLaunchKernelGGL((k_sddmm<PREC, L>), dim3(G, chunks), dim3(kBlock), 0, ctx->stream,         \
                       A->rowptr, A->colind, win, M, (int)beta[0]->NH, RW);                       \
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
                         A->rowptr, A->colind, win, M, (int)beta[0]->NH, RW);
  }
#undef KPM_SDDMM
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

int launch_sddmm(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* const* beta,
                 const kpm_block* const* alpha, const double* c, int W, double2* M) {
  if (W < 1 || W > kWin) return fail(KPM_EINVAL, "sddmm window out of range");
  return beta[0]->prec == KPM_VEC_F32 ? launch_sddmm_t<KPM_VEC_F32>(ctx, A, beta, alpha, c, W, M)
                                      : launch_sddmm_t<KPM_VEC_F64>(ctx, A, beta, alpha, c, W, M);
}

int adj_step(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* x, const kpm_block* z,
             kpm_block* out, const kpm_block* src, double alpha, double beta, double gamma, int mode) {
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
  if (ckpt_stride < 0) return fail(KPM_EINVAL, "ckpt_stride must be >= 0");
  if ((int64_t)H->csc_to_csr.size() != H->nnz)
    return fail(KPM_EINVAL, "operator was not built from CSC arrays (kpm_csr_from_csc)");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int64_t nalpha = NC - 1;  // alpha_0 .. alpha_{NC-2} are used
  const int64_t NRp = padded_nr(NR, vec_prec);
  const size_t bbytes = (size_t)(NH * NRp) * elem_bytes(vec_prec);
  const int64_t nbr = kWin + 2;   // beta ring: the recurrence's two plus a window
  // storage plan: every alpha when asked (stride 1) or when it fits in 60% of
  // the free memory plus the workspace this context already holds (stride 0);
  // checkpoint pairs every K steps otherwise
  size_t fr = 0, tot = 0;
  KPM_HIP_TRY(hipMemGetInfo(&fr, &tot));
  const size_t avail = fr + ctx->adjoint.bytes;
  const bool full = ckpt_stride == 1 ||
                    (ckpt_stride == 0 && bbytes * (size_t)(1 + nbr + nalpha) <= avail / 10 * 6);
  const int64_t K = full ? 1 : (ckpt_stride > 1 ? ckpt_stride : (int64_t)std::ceil(std::sqrt((double)NC)));
  const int64_t nseg = full ? 0 : (nalpha + K - 1) / K;
  // workspace: v | beta ring | (full) alpha_0..alpha_{nalpha-1}  or
  //                           (checkpoints) forward ring, 2*nseg pairs, K segment
  // It stays with the context (grow-only, released by kpm_ctx_destroy), so a
  // loop of gradient steps allocates it once.
  const int64_t nblk = full ? 1 + nbr + nalpha : 1 + nbr + 2 + 2 * nseg + K;
  Buffer& ws = ctx->adjoint;
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
  kpm_block v = blk(0);
  auto beta = [&](int64_t j) { return blk(1 + j % nbr); };
  auto alpha_full = [&](int64_t i) { return i == 0 ? v : blk(1 + nbr + i); };
  auto ckA = [&](int64_t s) { return blk(1 + nbr + 2 + 2 * s); };
  auto ckB = [&](int64_t s) { return blk(1 + nbr + 3 + 2 * s); };
  auto seg = [&](int64_t i) { return blk(1 + nbr + 2 + 2 * nseg + i); };
  Buffer& Mbuf = ctx->adj_grad;
  int rc = KPM_OK;
  do {
    const size_t mbytes = (size_t)std::max<int64_t>(H->nnz, 1) * sizeof(double2);
    if ((rc = ensure(ctx, Mbuf, mbytes)) != KPM_OK) break;
    KPM_HIP_TRY(hipMemsetAsync(Mbuf.ptr, 0, mbytes, ctx->stream));
    // v's padded columns and the beta ring (beta_NC = beta_{NC+1} = 0) start at zero
    KPM_HIP_TRY(hipMemsetAsync(ws.ptr, 0, bbytes * (size_t)(1 + nbr), ctx->stream));
    if ((rc = ensure(ctx, ctx->staging, (size_t)(NH * NR) * 16)) != KPM_OK) break;
    KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, psi, (size_t)(NH * NR) * 16, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = launch_upload(ctx, ctx->staging.ptr, &v)) != KPM_OK) break;

    // pending gradient terms (beta_j, alpha_{j-1}, c_j), added kWin at a time
    kpm_block pb[kWin], pa[kWin];
    const kpm_block* pbp[kWin];
    const kpm_block* pap[kWin];
    double pc[kWin];
    int np = 0;
    auto flush = [&]() -> int {
      if (np == 0) return KPM_OK;
      for (int w = 0; w < np; ++w) {
        pbp[w] = &pb[w];
        pap[w] = &pa[w];
      }
      const int r = launch_sddmm(ctx, H, pbp, pap, pc, np, (double2*)Mbuf.ptr);
      np = 0;
      return r;
    };
    // one reverse step: beta_j = g_j v + 2 H beta_{j+1} - beta_{j+2}, then queue
    // its term with alpha_{j-1}
    auto reverse = [&](int64_t j, const kpm_block& a_prev) -> int {
      kpm_block bx = beta(j + 1), bz = beta(j + 2), bo = beta(j);
      KPM_TRY(adj_step(ctx, H, &bx, &bz, &bo, &v, 2.0, -1.0, g[j], MODE_SRC));
      pb[np] = bo;
      pa[np] = a_prev;
      pc[np] = j == 1 ? 1.0 : 2.0;
      if (++np == kWin) return flush();
      return KPM_OK;
    };

    if (full) {
      // forward: alpha_1 = H v, alpha_k = 2 H alpha_{k-1} - alpha_{k-2}, each kept
      for (int64_t k = 1; k < nalpha && rc == KPM_OK; ++k) {
        kpm_block x = alpha_full(k - 1), o = alpha_full(k);
        if (k == 1) {
          rc = adj_step(ctx, H, &x, nullptr, &o, nullptr, 1.0, 0.0, 0.0, MODE_NONE);
        } else {
          kpm_block z = alpha_full(k - 2);
          rc = adj_step(ctx, H, &x, &z, &o, nullptr, 2.0, -1.0, 0.0, MODE_NONE);
        }
      }
      for (int64_t j = NC - 1; j >= 1 && rc == KPM_OK; --j) rc = reverse(j, alpha_full(j - 1));
      if (rc == KPM_OK) rc = flush();
    } else {
      // forward with checkpoint pairs (alpha_{sK}, alpha_{sK+1})
      kpm_block ring[2] = {blk(1 + nbr), blk(2 + nbr)};
      if ((rc = kpm_block_copy(&ring[0], &v)) != KPM_OK) break;
      if ((rc = adj_step(ctx, H, &ring[0], nullptr, &ring[1], nullptr, 1.0, 0.0, 0.0, MODE_NONE)) != KPM_OK) break;
      const int64_t klast = (nseg - 1) * K + 1;
      for (int64_t k = 1; k <= klast && rc == KPM_OK; ++k) {
        if (k >= 2)
          rc = adj_step(ctx, H, &ring[(k - 1) % 2], &ring[k % 2], &ring[k % 2], nullptr, 2.0, -1.0, 0.0, MODE_NONE);
        if (rc == KPM_OK && (k - 1) % K == 0) {  // pair (alpha_{k-1}, alpha_k) starts segment (k-1)/K
          const int64_t sg = (k - 1) / K;
          kpm_block a = ckA(sg), b = ckB(sg);
          rc = kpm_block_copy(&a, &ring[(k - 1) % 2]);
          if (rc == KPM_OK) rc = kpm_block_copy(&b, &ring[k % 2]);
        }
      }
      // reverse, one recomputed segment at a time; pending terms are flushed
      // before the next segment overwrites the segment buffers
      for (int64_t sg = nseg - 1; sg >= 0 && rc == KPM_OK; --sg) {
        const int64_t lo = sg * K, hi = std::min(lo + K - 1, nalpha - 1);
        kpm_block s0 = seg(0), a = ckA(sg), b = ckB(sg);
        rc = kpm_block_copy(&s0, &a);
        if (rc == KPM_OK && hi > lo) {
          kpm_block s1 = seg(1);
          rc = kpm_block_copy(&s1, &b);
        }
        for (int64_t i = lo + 2; i <= hi && rc == KPM_OK; ++i) {
          kpm_block si = seg(i - lo), sm1 = seg(i - lo - 1), sm2 = seg(i - lo - 2);
          rc = adj_step(ctx, H, &sm1, &sm2, &si, nullptr, 2.0, -1.0, 0.0, MODE_NONE);
        }
        for (int64_t i = hi; i >= lo && rc == KPM_OK; --i) rc = reverse(i + 1, seg(i - lo));
        if (rc == KPM_OK) rc = flush();
      }
    }
    if (rc != KPM_OK) break;
    std::vector<double> m(2 * std::max<int64_t>(H->nnz, 1));
    KPM_HIP_TRY(hipMemcpyAsync(m.data(), Mbuf.ptr, mbytes, hipMemcpyDeviceToHost, ctx->stream));
    KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const double inv = 1.0 / (double)NR;
    for (int64_t k = 0; k < H->nnz; ++k) {
      const int64_t p = H->csc_to_csr[k];
      grad[2 * k] = m[2 * p] * inv;
      grad[2 * k + 1] = m[2 * p + 1] * inv;
    }
  } while (false);
  (void)hipStreamSynchronize(ctx->stream);
  return rc;
}

// Host bookkeeping for the bond parametrisation H_ij = t_ij u_ij,
// H_ji = t_ij conj(u_ij): for every stored pair i < j of a CSC pattern that
// contains both (i, j) and (j, i), dL/dt_ij = Re(u M_ij + conj(u) M_ji) / a
// with u = H_ij / |H_ij| (1 when H_ij = 0).  O(nnz) with one merge per column.
int kpm_bond_gradient(int64_t n, const int64_t* colptr, const int64_t* rowval, const double* h,
                      const double* grad, double a, int64_t* bi, int64_t* bj, double* val,
                      int64_t* nbonds) {
  if (n <= 0 || !colptr || !rowval || !h || !grad || !bi || !bj || !val || !nbonds)
    return fail(KPM_EINVAL, "null argument");
  if (!(a > 0.0)) return fail(KPM_EINVAL, "a must be positive");
  // position of (r, c) for every stored entry: search column r for row c
  auto find = [&](int64_t col, int64_t row) -> int64_t {
    const int64_t* b = rowval + colptr[col];
    const int64_t* e = rowval + colptr[col + 1];
    const int64_t* p = std::lower_bound(b, e, row);
    return (p != e && *p == row) ? (int64_t)(p - rowval) : -1;
  };
  int64_t m = 0;
  for (int64_t c = 0; c < n; ++c)
    for (int64_t k = colptr[c]; k < colptr[c + 1]; ++k) {
      const int64_t r = rowval[k];
      if (r >= c) continue;  // entry (r, c) with r < c: pair i = r, j = c
      const int64_t kt = find(r, c);  // entry (c, r)
      if (kt < 0) return fail(KPM_EINVAL, "pattern is not symmetric: entry (j, i) missing");
      const double hr = h[2 * k], hi = h[2 * k + 1];
      const double mag = std::hypot(hr, hi);
      const double ur = mag > 0.0 ? hr / mag : 1.0, ui = mag > 0.0 ? hi / mag : 0.0;
      const double gr = grad[2 * k], gi = grad[2 * k + 1];
      const double tr = grad[2 * kt], ti = grad[2 * kt + 1];
      bi[m] = r;
      bj[m] = c;
      val[m] = (ur * (gr + tr) - ui * (gi - ti)) / a;
      ++m;
    }
  *nbonds = m;
  return KPM_OK;
}

}  // extern "C"
