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

#include "dot_steps.h"
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
// lanes (group_sum, a fixed order).  Rows wider than 64 words run as row chunks
// (gridDim.y > 1), each adding into its own accumulator M + chunk * mstride;
// the caller sums the chunks in order, so every entry is deterministic.
template <int PREC, int LPR>
__global__ void __launch_bounds__(kBlock)
    k_sddmm(const int* __restrict__ rowptr, const int* __restrict__ colind, SddmmWindow<PREC> win,
            double2* __restrict__ M, int NH, int RW, size_t mstride) {
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
      pr = group_sum<LPR>(pr);
      pi = group_sum<LPR>(pi);
      if (v && sub == 0) {
        double2* Mc = M + (size_t)blockIdx.y * mstride;
        double2 m = Mc[k];
        m.x += pr;
        m.y += pi;
        Mc[k] = m;
      }
    }
  }
}

// The same sums for rows of LPR >= 2 lanes, at most KB <= 16 nonzeros and a
// single row chunk, arranged for loads in flight: with G = min(LPR, 16),
// lane j of a row holds column index and accumulator entry of nonzeros
// j, j + G, ... (one coalesced load each); the window's beta words are
// loaded once and scaled by c_w (1 or 2: exact); then for each nonzero slot
// up to the wave's longest row, group_bcast hands the column to the row and
// the W gathers of alpha_w at that column go out together, are summed over
// the windows in order and reduced over the row's lanes (group_sum); each
// lane finally adds its own entries and stores them coalesced.  Slots past a
// row's end issue no load.
template <int PREC, int LPR, int KB>
__global__ void __launch_bounds__(kBlock)
    k_sddmm_rows(const int* __restrict__ rowptr, const int* __restrict__ colind,
                 SddmmWindow<PREC> win, double2* __restrict__ M, int NH, int RW) {
  using WT = Word<PREC>;
  using T = typename WT::T;
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
    T bw[kWin];
#pragma unroll
    for (int w = 0; w < kWin; ++w) bw[w] = w < win.W ? ld_stream(&win.beta[w][(size_t)rc * RW + sub]) : T{};
    const int s0 = rowptr[rc];
    const int n = live ? rowptr[rc + 1] - s0 : 0;
    int nmax = n;
#pragma unroll
    for (int o = LPR; o < 64; o <<= 1) nmax = max(nmax, __shfl_xor(nmax, o));
    int cl[ME];
    double2 m[ME];
#pragma unroll
    for (int e = 0; e < ME; ++e) {
      const int idx = e * MG + j;
      const bool v = idx < n;
      const int k = v ? s0 + idx : 0;
      cl[e] = v ? colind[k] : 0;
      m[e] = v ? ld_stream(&M[k]) : make_double2(0.0, 0.0);
    }
    double br[kWin][CPL], bi[kWin][CPL];
#pragma unroll
    for (int w = 0; w < kWin; ++w) {
      WT::unpack(bw[w], br[w], bi[w]);
#pragma unroll
      for (int q = 0; q < CPL; ++q) {
        br[w][q] *= win.c[w];
        bi[w][q] *= win.c[w];
      }
    }
    static_for<0, KB>([&](auto J) {
      constexpr int jj = decltype(J)::value;
      if (jj < nmax) {  // uniform
        const int col = group_bcast<MG, jj % MG>(cl[jj / MG]);
        const bool v = jj < n;
        const size_t off = (size_t)col * RW + sub;
        T av[kWin];
#pragma unroll
        for (int w = 0; w < kWin; ++w) av[w] = (w < win.W && v) ? win.alpha[w][off] : T{};
        double pr = 0.0, pi = 0.0;
#pragma unroll
        for (int w = 0; w < kWin; ++w) {
          double ar[CPL], ai[CPL];
          WT::unpack(av[w], ar, ai);
#pragma unroll
          for (int q = 0; q < CPL; ++q) {
            pr = fma(br[w][q], ar[q], fma(bi[w][q], ai[q], pr));
            pi = fma(br[w][q], ai[q], fma(-bi[w][q], ar[q], pi));
          }
        }
        pr = group_sum<LPR>(pr);
        pi = group_sum<LPR>(pi);
        if (j == jj % MG) {
          m[jj / MG].x += pr;
          m[jj / MG].y += pi;
        }
      }
    });
#pragma unroll
    for (int e = 0; e < ME; ++e) {
      const int idx = e * MG + j;
      if (owner && idx < n) st_stream(&M[s0 + idx], m[e]);
    }
  }
}

// out[k] = inv * sum_c M[perm[k] + c * mstride] (row chunks in order): the
// gradient of the k-th stored entry of the caller's CSC arrays.
__global__ void __launch_bounds__(kBlock)
    k_csc_gradient(const double2* __restrict__ M, const int* __restrict__ perm, int chunks,
                   size_t mstride, double inv, double2* __restrict__ out, int64_t nnz) {
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * kBlock) {
    const size_t p = (size_t)perm[k];
    double gr = 0.0, gi = 0.0;
    for (int c = 0; c < chunks; ++c) {
      const double2 m = M[p + c * mstride];
      gr += m.x;
      gi += m.y;
    }
    out[k] = make_double2(gr * inv, gi * inv);
  }
}

// resident workgroups per CU of the gradient-term launches (grid stride)
#ifndef KPM_SDDMM_WG_PER_CU
#define KPM_SDDMM_WG_PER_CU 8
#endif

template <int PREC>
static int launch_sddmm_t(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* const* beta,
                          const kpm_block* const* alpha, const double* c, int W, double2* M,
                          size_t mstride) {
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
  const unsigned G =
      (unsigned)std::min<int64_t>(nb, (int64_t)std::max(ctx->num_cus, 8) * KPM_SDDMM_WG_PER_CU);
  // the row kernel takes one block of rows per workgroup (a one-shot grid,
  // handed out in row order: 131.3 against 134.4 ms per c5 gradient step for
  // the persistent grid-stride form)
  const unsigned Grows = (unsigned)std::min<int64_t>(nb, (int64_t)INT32_MAX);
  const int w = A->max_row_nnz;
  if (lpr >= 2 && chunks == 1 && w <= 16) {
    const int kb = w <= 4 ? 4 : (w <= 8 ? 8 : (w <= 12 ? 12 : 16));
#define KPM_SDDMM_ROWS(L, KBv)                                                                    \
  if (lpr == L && kb == KBv) {                                                                    \
    hipLaunchKernelGGL((k_sddmm_rows<PREC, L, KBv>), dim3(Grows), dim3(kBlock), 0, ctx->stream,   \
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
                       A->rowptr, A->colind, win, M, (int)beta[0]->NH, RW, mstride);              \
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
                         A->rowptr, A->colind, win, M, (int)beta[0]->NH, RW, mstride);
  }
#undef KPM_SDDMM
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

int sddmm_chunks(int64_t NR, int prec) {
  const int RW = row_words(NR, prec);
  return RW > 64 ? RW / 64 : 1;
}

int launch_sddmm(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* const* beta,
                 const kpm_block* const* alpha, const double* c, int W, double2* M,
                 size_t mstride) {
  if (W < 1 || W > kWin) return fail(KPM_EINVAL, "sddmm window out of range");
  if (sddmm_chunks(beta[0]->NR, beta[0]->prec) > 1 && mstride < (size_t)A->nnz)
    return fail(KPM_EINVAL, "sddmm: row chunks need one accumulator each");
  return beta[0]->prec == KPM_VEC_F32
             ? launch_sddmm_t<KPM_VEC_F32>(ctx, A, beta, alpha, c, W, M, mstride)
             : launch_sddmm_t<KPM_VEC_F64>(ctx, A, beta, alpha, c, W, M, mstride);
}

int adj_step(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* x, const kpm_block* z,
             kpm_block* out, const kpm_block* src, double alpha, double beta, double gamma, int mode) {
  StepLaunch s{A, x, z, out, src, alpha, beta, mode, gamma};
  return launch_step(ctx, s, tuned_config(ctx, A, x->NR, x->prec, MODE_NONE), nullptr);
}

}  // namespace kpm

namespace kpm {

// Workspace of the split adjoint: v | beta ring (nbr) | then either every
// alpha_1..alpha_{nalpha-1} (full) or the forward ring (2), 2 * nseg
// checkpoint blocks and one segment of K blocks.
struct AdjLayout {
  int64_t NC, NR, NRp, NH, nalpha, nbr, K, nseg;
  int prec;
  bool full;
  size_t bbytes;
  int64_t nblk() const { return full ? 1 + nbr + nalpha : 1 + nbr + 2 + 2 * nseg + K; }
};

static AdjLayout adj_layout(kpm_ctx* ctx, int64_t NC, int64_t NR, int64_t NH, int prec,
                            int64_t ckpt_stride, size_t avail) {
  AdjLayout L;
  L.NC = NC;
  L.NR = NR;
  L.NH = NH;
  L.prec = prec;
  L.NRp = padded_nr(NR, prec);
  L.bbytes = (size_t)(NH * L.NRp) * elem_bytes(prec);
  L.nalpha = NC - 1;  // alpha_0 .. alpha_{NC-2} enter the gradient
  L.nbr = kWin + 2;   // beta ring: the recurrence's two plus a window
  L.full = ckpt_stride == 1 ||
           (ckpt_stride == 0 && L.bbytes * (size_t)(1 + L.nbr + L.nalpha) <= avail / 10 * 6);
  L.K = L.full ? 1 : (ckpt_stride > 1 ? ckpt_stride : (int64_t)std::ceil(std::sqrt((double)NC)));
  L.nseg = L.full ? 0 : (L.nalpha + L.K - 1) / L.K;
  (void)ctx;
  return L;
}

static kpm_block adj_blk(kpm_ctx* ctx, const AdjLayout& L, int64_t i) {
  kpm_block b;
  b.ctx = ctx;
  b.NH = L.NH;
  b.NR = L.NR;
  b.NRp = L.NRp;
  b.prec = L.prec;
  b.d = (char*)ctx->adjoint.ptr + L.bbytes * (size_t)i;
  b.owned = false;
  return b;
}

// forward pass: keeps alpha (or checkpoint pairs) and writes the moments
// mu_n = <v|alpha_n> (n < NC, device NR x NC) from the same passes.  The
// moments come from the doubling identities of kpm_1d! (src/moment.jl:442-487):
// pass k <= ceil(NC/2) dots <alpha_{k-1}|alpha_{k-1}> and
// <alpha_{k-1}|alpha_k>, which give mu_{2k-2} and mu_{2k-1}; later passes
// carry no dots.  No pass reads more than the recurrence's own vectors.
static int adj_forward(kpm_ctx* ctx, const kpm_csr* H, const AdjLayout& L, const double* psi) {
  KPM_TRY(ensure(ctx, ctx->adjoint, L.bbytes * (size_t)L.nblk()));
  kpm_block v = adj_blk(ctx, L, 0);
  KPM_HIP_TRY(hipMemsetAsync(v.d, 0, L.bbytes, ctx->stream));
  KPM_TRY(ensure(ctx, ctx->staging, (size_t)(L.NH * L.NR) * 16));
  KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, psi, (size_t)(L.NH * L.NR) * 16,
                             hipMemcpyHostToDevice, ctx->stream));
  KPM_TRY(launch_upload(ctx, ctx->staging.ptr, &v));
  // one spare column: for odd NC the last dot pass also yields mu_NC
  KPM_TRY(ensure(ctx, ctx->mu, (size_t)(L.NR * (L.NC + 1)) * sizeof(double2)));
  const int64_t ndot = (L.NC + 1) / 2;  // passes that carry dots
  DotSteps ds{ctx, H, L.NR, L.NRp, {nullptr, nullptr}};
  KPM_TRY(ds.init(2, &v, MODE_DBL));
  auto dbl = [](int64_t k) {
    if (k == 1) return Combine{2, {0, 1}, {1.0, 1.0}, {-1, -1}, {0.0, 0.0}};
    return Combine{2, {(int)(2 * k - 2), (int)(2 * k - 1)}, {2.0, 2.0}, {0, 1}, {-1.0, -1.0}};
  };
  // pass k: out = alpha_k from x = alpha_{k-1}, z = alpha_{k-2}
  auto pass = [&](int64_t k, kpm_block* x, kpm_block* z, kpm_block* o) -> int {
    const double a = k == 1 ? 1.0 : 2.0, b = k == 1 ? 0.0 : -1.0;
    if (k <= ndot) {
      KPM_TRY(ds.step(x, k == 1 ? nullptr : z, o, nullptr, a, b, MODE_DBL, dbl(k)));
      if (k == ndot) KPM_TRY(ds.finish());
      return KPM_OK;
    }
    return adj_step(ctx, H, x, k == 1 ? nullptr : z, o, nullptr, a, b, 0.0, MODE_NONE);
  };
  kpm_block scratch = adj_blk(ctx, L, 1);  // a beta-ring block, cleared by the reverse pass
  if (L.full) {
    auto alpha = [&](int64_t i) { return i == 0 ? v : adj_blk(ctx, L, 1 + L.nbr + i); };
    for (int64_t k = 1; k < L.NC; ++k) {
      kpm_block x = alpha(k - 1), z = k >= 2 ? alpha(k - 2) : v, o = k < L.nalpha ? alpha(k) : scratch;
      KPM_TRY(pass(k, &x, &z, &o));
    }
  } else {
    // a two-block ring with checkpoint pairs (alpha_{sK}, alpha_{sK+1})
    kpm_block ring[2] = {adj_blk(ctx, L, 1 + L.nbr), adj_blk(ctx, L, 2 + L.nbr)};
    KPM_TRY(kpm_block_copy(&ring[0], &v));
    auto ckA = [&](int64_t sg) { return adj_blk(ctx, L, 1 + L.nbr + 2 + 2 * sg); };
    auto ckB = [&](int64_t sg) { return adj_blk(ctx, L, 1 + L.nbr + 3 + 2 * sg); };
    for (int64_t k = 1; k < L.NC; ++k) {
      // alpha_k overwrites alpha_{k-2} in place
      KPM_TRY(pass(k, &ring[(k - 1) % 2], &ring[k % 2], &ring[k % 2]));
      if ((k - 1) % L.K == 0 && (k - 1) / L.K < L.nseg) {  // pair starting segment (k-1)/K
        const int64_t sg = (k - 1) / L.K;
        kpm_block a = ckA(sg), b = ckB(sg);
        KPM_TRY(kpm_block_copy(&a, &ring[(k - 1) % 2]));
        KPM_TRY(kpm_block_copy(&b, &ring[k % 2]));
      }
    }
  }
  return KPM_OK;
}

// reverse pass from g = dL/dmu on the kept forward vectors; grad in the CSC
// order of H (scaled by 1/NR)
static int adj_reverse(kpm_ctx* ctx, const kpm_csr* H, const AdjLayout& L, const double* g,
                       double* grad) {
  kpm_block v = adj_blk(ctx, L, 0);
  auto beta = [&](int64_t j) { return adj_blk(ctx, L, 1 + j % L.nbr); };
  auto alpha_full = [&](int64_t i) { return i == 0 ? v : adj_blk(ctx, L, 1 + L.nbr + i); };
  auto ckA = [&](int64_t sg) { return adj_blk(ctx, L, 1 + L.nbr + 2 + 2 * sg); };
  auto ckB = [&](int64_t sg) { return adj_blk(ctx, L, 1 + L.nbr + 3 + 2 * sg); };
  auto seg = [&](int64_t i) { return adj_blk(ctx, L, 1 + L.nbr + 2 + 2 * L.nseg + i); };
  const int chunks = sddmm_chunks(L.NR, L.prec);
  const size_t nnz = (size_t)std::max<int64_t>(H->nnz, 1);
  const size_t mbytes = nnz * chunks * sizeof(double2);
  KPM_TRY(ensure(ctx, ctx->adj_grad, mbytes));
  double2* M = (double2*)ctx->adj_grad.ptr;
  KPM_HIP_TRY(hipMemsetAsync(M, 0, mbytes, ctx->stream));
  // beta_NC = beta_{NC+1} = 0
  KPM_HIP_TRY(hipMemsetAsync(beta(0).d, 0, L.bbytes * (size_t)L.nbr, ctx->stream));

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
    const int r = launch_sddmm(ctx, H, pbp, pap, pc, np, M, nnz);
    np = 0;
    return r;
  };
  // beta_j = g_j v + 2 H beta_{j+1} - beta_{j+2}, then queue its term with alpha_{j-1}
  auto reverse = [&](int64_t j, const kpm_block& a_prev) -> int {
    kpm_block bx = beta(j + 1), bz = beta(j + 2), bo = beta(j);
    KPM_TRY(adj_step(ctx, H, &bx, &bz, &bo, &v, 2.0, -1.0, g[j], MODE_SRC));
    pb[np] = bo;
    pa[np] = a_prev;
    pc[np] = j == 1 ? 1.0 : 2.0;
    if (++np == kWin) return flush();
    return KPM_OK;
  };
  if (L.full) {
    for (int64_t j = L.NC - 1; j >= 1; --j) KPM_TRY(reverse(j, alpha_full(j - 1)));
    KPM_TRY(flush());
  } else {
    // one recomputed segment at a time; pending terms are flushed before the
    // next segment overwrites the segment buffers
    for (int64_t sg = L.nseg - 1; sg >= 0; --sg) {
      const int64_t lo = sg * L.K, hi = std::min(lo + L.K - 1, L.nalpha - 1);
      kpm_block s0 = seg(0), a = ckA(sg), b = ckB(sg);
      KPM_TRY(kpm_block_copy(&s0, &a));
      if (hi > lo) {
        kpm_block s1 = seg(1);
        KPM_TRY(kpm_block_copy(&s1, &b));
      }
      for (int64_t i = lo + 2; i <= hi; ++i) {
        kpm_block si = seg(i - lo), sm1 = seg(i - lo - 1), sm2 = seg(i - lo - 2);
        KPM_TRY(adj_step(ctx, H, &sm1, &sm2, &si, nullptr, 2.0, -1.0, 0.0, MODE_NONE));
      }
      for (int64_t i = hi; i >= lo; --i) KPM_TRY(reverse(i + 1, seg(i - lo)));
      KPM_TRY(flush());
    }
  }
  // CSC order on the device, then one download into the caller's array
  if (H->nnz > 0) {
    if (!H->csc_to_csr_dev) {
      int* d = nullptr;
      KPM_HIP_TRY(hipMalloc(&d, (size_t)H->nnz * sizeof(int)));
      hipError_t e = hipMemcpyAsync(d, H->csc_to_csr.data(), (size_t)H->nnz * sizeof(int),
                                    hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) {
        (void)hipFree(d);
        return fail(KPM_EHIP, std::string("gradient permutation upload: ") + hipGetErrorString(e));
      }
      const_cast<kpm_csr*>(H)->csc_to_csr_dev = d;
    }
    KPM_TRY(ensure(ctx, ctx->staging, (size_t)H->nnz * sizeof(double2)));
    const int64_t gb = std::min<int64_t>((H->nnz + kBlock - 1) / kBlock, 65535);
    hipLaunchKernelGGL(k_csc_gradient, dim3((unsigned)gb), dim3(kBlock), 0, ctx->stream, M,
                       H->csc_to_csr_dev, chunks, nnz, 1.0 / (double)L.NR,
                       (double2*)ctx->staging.ptr, H->nnz);
    KPM_HIP_TRY(hipGetLastError());
    KPM_HIP_TRY(hipMemcpyAsync(grad, ctx->staging.ptr, (size_t)H->nnz * sizeof(double2),
                               hipMemcpyDeviceToHost, ctx->stream));
  }
  ++ctx->counters[KPM_COUNT_SYNC];
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return KPM_OK;
}

static int adj_check(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t NR, int64_t NH, int vec_prec,
                     int64_t ckpt_stride) {
  if (!ctx || !H) return fail(KPM_EINVAL, "null argument");
  if (NC < 2 || NR <= 0 || NH != H->n) return fail(KPM_EINVAL, "bad sizes");
  if (vec_prec != KPM_VEC_F64 && vec_prec != KPM_VEC_F32) return fail(KPM_EINVAL, "bad vec_prec");
  if (ckpt_stride < 0) return fail(KPM_EINVAL, "ckpt_stride must be >= 0");
  if ((int64_t)H->csc_to_csr.size() != H->nnz)
    return fail(KPM_EINVAL, "operator was not built from CSC arrays (kpm_csr_from_csc)");
  return KPM_OK;
}

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_1d_adjoint_forward(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t NR, int64_t NH,
                           const double* psi, double* mu_all, int vec_prec, int64_t ckpt_stride) {
  KPM_TRY(adj_check(ctx, H, NC, NR, NH, vec_prec, ckpt_stride));
  if (!psi || !mu_all) return fail(KPM_EINVAL, "null argument");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  ctx->adj1.valid = false;
  // storage plan: every alpha when asked (stride 1) or when it fits in 60% of
  // the free memory plus the workspace this context already holds (stride
  // 0); checkpoint pairs every K steps otherwise.  The workspace stays with
  // the context (grow-only, released by kpm_ctx_destroy).
  size_t fr = 0, tot = 0;
  KPM_HIP_TRY(hipMemGetInfo(&fr, &tot));
  const AdjLayout L = adj_layout(ctx, NC, NR, NH, vec_prec, ckpt_stride, fr + ctx->adjoint.bytes);
  int rc = adj_forward(ctx, H, L, psi);
  if (rc == KPM_OK) {
    hipError_t e = hipMemcpyAsync(mu_all, ctx->mu.ptr, (size_t)(NR * NC) * sizeof(double2),
                                  hipMemcpyDeviceToHost, ctx->stream);
    ++ctx->counters[KPM_COUNT_SYNC];
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = fail(KPM_EHIP, std::string("moments download: ") + hipGetErrorString(e));
  }
  if (rc != KPM_OK) {
    (void)hipStreamSynchronize(ctx->stream);
    return rc;
  }
  ctx->adj1 = kpm_adj1_state{true, H->serial, NC, NR, NH, vec_prec, ckpt_stride,
                             L.full ? 1 : L.K};
  return KPM_OK;
}

int kpm_1d_adjoint_reverse(kpm_ctx* ctx, const kpm_csr* H, const double* g, double* grad) {
  if (!ctx || !H || !g || !grad) return fail(KPM_EINVAL, "null argument");
  const kpm_adj1_state& S = ctx->adj1;
  if (!S.valid || S.H_serial != H->serial)
    return fail(KPM_EINVAL, "no kept forward pass of this operator (kpm_1d_adjoint_forward)");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  // the layout the forward pass chose (its stride decision is kept)
  const AdjLayout L = adj_layout(ctx, S.NC, S.NR, S.NH, S.prec, S.stride_used, 0);
  const int rc = adj_reverse(ctx, H, L, g, grad);
  (void)hipStreamSynchronize(ctx->stream);
  return rc;
}

int kpm_1d_adjoint(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, int64_t NR, int64_t NH,
                   const double* psi, const double* g, double* grad, int vec_prec,
                   int64_t ckpt_stride) {
  if (!g || !grad) return fail(KPM_EINVAL, "null argument");
  std::vector<double> mu((size_t)(2 * std::max<int64_t>(NR, 0) * std::max<int64_t>(NC, 0)));
  KPM_TRY(kpm_1d_adjoint_forward(ctx, H, NC, NR, NH, psi, mu.data(), vec_prec, ckpt_stride));
  return kpm_1d_adjoint_reverse(ctx, H, g, grad);
}

// Host bookkeeping for the bond parametrisation H_ij = t_ij u_ij,
// H_ji = t_ij conj(u_ij): for every stored pair i < j of a CSC pattern that
// contains both (i, j) and (j, i), dL/dt_ij = Re(u M_ij + conj(u) M_ji) / a
// with u = H_ij / |H_ij| (1 when H_ij = 0).  O(nnz); columns sorted by row.
int kpm_bond_gradient(int64_t n, const int64_t* colptr, const int64_t* rowval, const double* h,
                      const double* grad, double a, int64_t* bi, int64_t* bj, double* val,
                      int64_t* nbonds) {
  if (n <= 0 || !colptr || !rowval || !h || !grad || !bi || !bj || !val || !nbonds)
    return fail(KPM_EINVAL, "null argument");
  if (!(a > 0.0)) return fail(KPM_EINVAL, "a must be positive");
  // Position of the transposed entry (c, r) of every stored (r, c): visiting
  // the columns c in increasing order, the entries (c, r) of column r come up
  // in increasing c, i.e. in column r's sorted order, so a cursor per column
  // walks them (one O(nnz) pass, no search).
  std::vector<int64_t> cur(colptr, colptr + n);
  int64_t m = 0;
  for (int64_t c = 0; c < n; ++c)
    for (int64_t k = colptr[c]; k < colptr[c + 1]; ++k) {
      const int64_t r = rowval[k];
      if (r < 0 || r >= n) return fail(KPM_EINVAL, "row index out of range");
      const int64_t kt = cur[r]++;  // entry (c, r)
      if (kt >= colptr[r + 1] || rowval[kt] != c)
        return fail(KPM_EINVAL,
                    "pattern is not symmetric (or a column is unsorted): entry (j, i) missing");
      if (r >= c) continue;  // entry (r, c) with r < c: pair i = r, j = c
      const double hr = h[2 * k], hi = h[2 * k + 1];
      const double mag = std::sqrt(hr * hr + hi * hi);  // |H_ij| of a Hamiltonian entry
      const double gr = grad[2 * k], gi = grad[2 * k + 1];
      const double tr = grad[2 * kt], ti = grad[2 * kt + 1];
      bi[m] = r;
      bj[m] = c;
      // Re(u (g + t) ...) with u = h / |h|, one division per bond
      val[m] = mag > 0.0 ? (hr * (gr + tr) - hi * (gi - ti)) / (mag * a) : (gr + tr) / a;
      ++m;
    }
  *nbonds = m;
  return KPM_OK;
}

}  // extern "C"
