// adjoint2d.hip -- reverse-mode gradient of the kpm_2d moment table.
//
// No counterpart in the reference.  The table of _kpm_2d_blocked!
// (src/moment.jl:987-1088) is, for n, m < NC,
//
//   mu[n,m] = (1/NR) sum_r <l_m | Jb r_n>,  l_m = T_m(H) psi_l,  r_n = T_n(H) Ja psi_r
//
// (column r of every block; <x|y> = x^H y).  For a real loss L with
// cotangent Gb (dL = Re sum_nm conj(Gb[n,m]) dmu[n,m]) the derivatives on the
// stored entries of H, Ja and Jb (dL = Re sum_k grad[k] dA_k) are
//
//   Jb:  M_ij = (1/NR) sum_m conj(l_m[i]) Q_m[j],   Q_m = sum_n conj(Gb[n,m]) r_n
//   Ja:  M_ij = (1/NR) conj(Y[i]) psi_r[j],         Y = sum_n T_n(H) z_n
//   H:   M_ij = (1/NR) sum_k c_k (conj(lam_k[i]) r_{k-1}[j] + conj(kap_k[i]) l_{k-1}[j])
//
// with P_n = sum_m Gb[n,m] l_m, z_n = Jb^H P_n, y_m = Jb Q_m, c_1 = 1,
// c_k = 2 (k >= 2), and the reverse Clenshaw vectors
//   lam_k = z_k + 2 H lam_{k+1} - lam_{k+2},  kap_k = y_k + 2 H kap_{k+1} - kap_{k+2}
// (lam_NC = lam_{NC+1} = 0; Y = z_0 + H lam_1 - lam_2).  The right chain's
// terms follow from d r_n = dT_n(H) a + T_n(H) dJa psi_r (a = Ja psi_r), the
// left chain's from d l_m = dT_m(H) psi_l, and each dT_k(H) v is the
// recurrence driven by dH T_{k-1} v, whose adjoint is the Clenshaw sum above.
//
// Every chain vector is kept: l_0..l_{NC-1}, r_0..r_{NC-1}, P and Q, i.e.
// 4 NC recurrence blocks plus rings (69 GB at NH = 131,072, NR = 16,
// NC = 512, FP64).  P and Q are complex GEMMs of the slots by an NC x NC
// matrix (k_mix); the sums over the pattern reuse the windowed SDDMM of the
// kpm_1d adjoint.
#include <algorithm>
#include <vector>

#include "step_kernel.h"

namespace kpm {

// out_j[k] = sum_{i < n_in} C[i + ldc j] A_i[k] for k < K, j < n_out: slot i
// of A starts at A + i * pitch_a, slot j of out at out + j * pitch_o
// (complex values).  A workgroup owns kBlock consecutive k and kMixJ
// consecutive j; the coefficients of 32 inputs at a time are staged in LDS
// and read as broadcasts.  Each sum runs over i in increasing order.
constexpr int kMixJ = 32;
constexpr int kMixI = 32;

__global__ void __launch_bounds__(kBlock)
    k_mix(const double2* __restrict__ A, int64_t pitch_a, int n_in, const double2* __restrict__ C,
          int ldc, double2* __restrict__ out, int64_t pitch_o, int n_out, int64_t K) {
  __shared__ double2 cs[kMixI][kMixJ];
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int j0 = blockIdx.y * kMixJ;
  const bool live = k < K;
  const int64_t kc = live ? k : 0;
  double accr[kMixJ], acci[kMixJ];
#pragma unroll
  for (int jj = 0; jj < kMixJ; ++jj) accr[jj] = acci[jj] = 0.0;
  for (int i0 = 0; i0 < n_in; i0 += kMixI) {
    __syncthreads();
    for (int t = threadIdx.x; t < kMixI * kMixJ; t += kBlock) {
      const int ii = t / kMixJ, jj = t % kMixJ;
      const int i = i0 + ii, j = j0 + jj;
      cs[ii][jj] = (i < n_in && j < n_out) ? C[i + (int64_t)ldc * j] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int ni = min(kMixI, n_in - i0);
    for (int ii = 0; ii < ni; ++ii) {
      const double2 a = A[(int64_t)(i0 + ii) * pitch_a + kc];
#pragma unroll
      for (int jj = 0; jj < kMixJ; ++jj) {
        const double2 c = cs[ii][jj];
        accr[jj] = fma(c.x, a.x, fma(-c.y, a.y, accr[jj]));
        acci[jj] = fma(c.x, a.y, fma(c.y, a.x, acci[jj]));
      }
    }
  }
  if (!live) return;
#pragma unroll
  for (int jj = 0; jj < kMixJ; ++jj)
    if (j0 + jj < n_out) out[(int64_t)(j0 + jj) * pitch_o + k] = make_double2(accr[jj], acci[jj]);
}

static int launch_mix(kpm_ctx* ctx, const void* A, int64_t pitch_a, int n_in, const void* C,
                      int ldc, void* out, int64_t pitch_o, int n_out, int64_t K) {
  const int64_t gx = (K + kBlock - 1) / kBlock;
  const int64_t gy = (n_out + kMixJ - 1) / kMixJ;
  if (gx > INT32_MAX || gy > 65535) return fail(KPM_EINVAL, "slot mix too large for one launch");
  hipLaunchKernelGGL(k_mix, dim3((unsigned)gx, (unsigned)gy), dim3(kBlock), 0, ctx->stream,
                     (const double2*)A, pitch_a, n_in, (const double2*)C, ldc, (double2*)out,
                     pitch_o, n_out, K);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_2d_adjoint(kpm_ctx* ctx, const kpm_csr* H, const kpm_csr* Ja, const kpm_csr* Jb,
                   const kpm_csr* Jb_adj, int64_t NC, int64_t NR, int64_t NH, const double* psi_l,
                   const double* psi_r, const double* gbar, double* grad_H, double* grad_Ja,
                   double* grad_Jb) {
  if (!ctx || !H || !Ja || !Jb || !Jb_adj || !psi_l || !psi_r || !gbar || !grad_H || !grad_Ja ||
      !grad_Jb)
    return fail(KPM_EINVAL, "null argument");
  if (NC < 1 || NR <= 0 || NH != H->n || Ja->n != NH || Jb->n != NH || Jb_adj->n != NH)
    return fail(KPM_EINVAL, "bad sizes");
  for (const kpm_csr* A : {H, Ja, Jb})
    if ((int64_t)A->csc_to_csr.size() != A->nnz)
      return fail(KPM_EINVAL, "H, Ja and Jb must come from kpm_csr_from_csc");
  if (NC > 65535 * kMixJ) return fail(KPM_EINVAL, "NC too large");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const int prec = KPM_VEC_F64;
  const int64_t NRp = padded_nr(NR, prec);
  const int64_t Kc = NH * NRp;  // complex values per block
  const size_t bbytes = (size_t)Kc * sizeof(double2);
  const int64_t nbr = kWin + 2;
  // blocks: L[NC] | R[NC] | P[NC] | Q[NC] | lam ring[nbr] | psi_r | tmp | Y
  const int64_t nblk = 4 * NC + nbr + 3;
  KPM_TRY(ensure(ctx, ctx->adj2, bbytes * (size_t)nblk));
  auto blk = [&](int64_t i) {
    kpm_block b;
    b.ctx = ctx;
    b.NH = NH;
    b.NR = NR;
    b.NRp = NRp;
    b.prec = prec;
    b.d = (char*)ctx->adj2.ptr + bbytes * (size_t)i;
    b.owned = false;
    return b;
  };
  auto Ls = [&](int64_t m) { return blk(m); };
  auto Rs = [&](int64_t n) { return blk(NC + n); };
  auto Ps = [&](int64_t n) { return blk(2 * NC + n); };
  auto Qs = [&](int64_t m) { return blk(3 * NC + m); };
  auto lam = [&](int64_t k) { return blk(4 * NC + k % nbr); };
  kpm_block pr = blk(4 * NC + nbr), tmp = blk(4 * NC + nbr + 1), Y = blk(4 * NC + nbr + 2);
  // gradient accumulators (CSR order) and the two coefficient matrices
  const size_t nH = (size_t)std::max<int64_t>(H->nnz, 1), nA = (size_t)std::max<int64_t>(Ja->nnz, 1),
               nB = (size_t)std::max<int64_t>(Jb->nnz, 1);
  // rows wider than 64 words: one accumulator per row chunk, summed in order
  const int chunks = sddmm_chunks(NR, prec);
  const size_t gbytes = (nH + nA + nB) * chunks * sizeof(double2);
  const size_t cbytes = (size_t)(2 * NC * NC) * sizeof(double2);
  KPM_TRY(ensure(ctx, ctx->adj2_aux, gbytes + cbytes));
  double2* MH = (double2*)ctx->adj2_aux.ptr;
  double2* MA = MH + nH * chunks;
  double2* MB = MA + nA * chunks;
  double2* CP = MB + nB * chunks;  // CP[m + NC n] = Gb[n, m]: P_n = sum_m CP[m, n] l_m
  double2* CQ = CP + NC * NC;      // CQ[n + NC m] = conj(Gb[n, m]): Q_m = sum_n CQ[n, m] r_n
  {
    std::vector<double> h(4 * (size_t)(NC * NC));
    for (int64_t m = 0; m < NC; ++m)
      for (int64_t n = 0; n < NC; ++n) {
        const double gr = gbar[2 * (n + NC * m)], gi = gbar[2 * (n + NC * m) + 1];
        h[2 * (m + NC * n)] = gr;
        h[2 * (m + NC * n) + 1] = gi;
        h[2 * (size_t)(NC * NC) + 2 * (n + NC * m)] = gr;
        h[2 * (size_t)(NC * NC) + 2 * (n + NC * m) + 1] = -gi;
      }
    KPM_HIP_TRY(hipMemcpyAsync(CP, h.data(), cbytes, hipMemcpyHostToDevice, ctx->stream));
    KPM_HIP_TRY(hipMemsetAsync(MH, 0, gbytes, ctx->stream));
    // the lam ring starts at zero (lam_NC = lam_{NC+1} = 0); padded probe
    // columns are zero in every uploaded block and stay zero
    KPM_HIP_TRY(hipMemsetAsync(lam(0).d, 0, bbytes * (size_t)nbr, ctx->stream));
    KPM_TRY(ensure(ctx, ctx->staging, (size_t)(NH * NR) * 16));
    kpm_block l0 = Ls(0);
    KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, psi_l, (size_t)(NH * NR) * 16, hipMemcpyHostToDevice,
                               ctx->stream));
    KPM_TRY(launch_upload(ctx, ctx->staging.ptr, &l0));
    KPM_HIP_TRY(hipMemcpyAsync(ctx->staging.ptr, psi_r, (size_t)(NH * NR) * 16, hipMemcpyHostToDevice,
                               ctx->stream));
    KPM_TRY(launch_upload(ctx, ctx->staging.ptr, &pr));
    KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));  // the host matrix h goes out of scope
  }
  // forward chains: l_m = T_m(H) psi_l, r_n = T_n(H) Ja psi_r (src/moment.jl:1013-1050)
  {
    kpm_block r0 = Rs(0);
    KPM_TRY(adj_step(ctx, Ja, &pr, nullptr, &r0, nullptr, 1.0, 0.0, 0.0, MODE_NONE));
    for (int chain = 0; chain < 2; ++chain)
      for (int64_t k = 1; k < NC; ++k) {
        kpm_block x = chain ? Rs(k - 1) : Ls(k - 1), o = chain ? Rs(k) : Ls(k);
        if (k == 1) {
          KPM_TRY(adj_step(ctx, H, &x, nullptr, &o, nullptr, 1.0, 0.0, 0.0, MODE_NONE));
        } else {
          kpm_block z = chain ? Rs(k - 2) : Ls(k - 2);
          KPM_TRY(adj_step(ctx, H, &x, &z, &o, nullptr, 2.0, -1.0, 0.0, MODE_NONE));
        }
      }
  }
  // P = L Gb^T and Q = R conj(Gb) (slot mixes)
  KPM_TRY(launch_mix(ctx, Ls(0).d, Kc, (int)NC, CP, (int)NC, Ps(0).d, Kc, (int)NC, Kc));
  KPM_TRY(launch_mix(ctx, Rs(0).d, Kc, (int)NC, CQ, (int)NC, Qs(0).d, Kc, (int)NC, Kc));

  // windowed SDDMM terms
  kpm_block wb[kWin], wa[kWin];
  const kpm_block* wbp[kWin];
  const kpm_block* wap[kWin];
  double wc[kWin];
  int nw = 0;
  auto flush = [&](const kpm_csr* A, double2* M) -> int {
    if (nw == 0) return KPM_OK;
    for (int w = 0; w < nw; ++w) {
      wbp[w] = &wb[w];
      wap[w] = &wa[w];
    }
    const int r = launch_sddmm(ctx, A, wbp, wap, wc, nw, M, (size_t)std::max<int64_t>(A->nnz, 1));
    nw = 0;
    return r;
  };
  auto queue = [&](const kpm_csr* A, double2* M, const kpm_block& b, const kpm_block& a,
                   double c) -> int {
    wb[nw] = b;
    wa[nw] = a;
    wc[nw] = c;
    if (++nw == kWin) return flush(A, M);
    return KPM_OK;
  };
  // Jb: sum_m conj(l_m) Q_m
  for (int64_t m = 0; m < NC; ++m) KPM_TRY(queue(Jb, MB, Ls(m), Qs(m), 1.0));
  KPM_TRY(flush(Jb, MB));

  // reverse Clenshaw over one chain: sources src(k) for k = NC-1 .. 1, terms
  // c_k conj(lam_k) fwd_{k-1} on H's pattern
  auto reverse = [&](auto src, auto fwd) -> int {
    KPM_HIP_TRY(hipMemsetAsync(lam(0).d, 0, bbytes * (size_t)nbr, ctx->stream));
    for (int64_t k = NC - 1; k >= 1; --k) {
      KPM_TRY(src(k));
      kpm_block bx = lam(k + 1), bz = lam(k + 2), bo = lam(k);
      KPM_TRY(adj_step(ctx, H, &bx, &bz, &bo, &tmp, 2.0, -1.0, 1.0, MODE_SRC));
      KPM_TRY(queue(H, MH, bo, fwd(k - 1), k == 1 ? 1.0 : 2.0));
    }
    return flush(H, MH);
  };
  // right chain: z_k = Jb^H P_k
  auto zsrc = [&](int64_t k) {
    kpm_block p = Ps(k);
    return adj_step(ctx, Jb_adj, &p, nullptr, &tmp, nullptr, 1.0, 0.0, 0.0, MODE_NONE);
  };
  KPM_TRY(reverse(zsrc, Rs));
  // Y = z_0 + H lam_1 - lam_2, then Ja: conj(Y) psi_r
  KPM_TRY(zsrc(0));
  if (NC >= 2) {
    kpm_block l1 = lam(1), l2 = lam(2);
    KPM_TRY(adj_step(ctx, H, &l1, &l2, &Y, &tmp, 1.0, -1.0, 1.0, MODE_SRC));
  } else {
    KPM_TRY(kpm_block_copy(&Y, &tmp));
  }
  KPM_TRY(queue(Ja, MA, Y, pr, 1.0));
  KPM_TRY(flush(Ja, MA));
  // left chain: y_k = Jb Q_k
  auto ysrc = [&](int64_t k) {
    kpm_block q = Qs(k);
    return adj_step(ctx, Jb, &q, nullptr, &tmp, nullptr, 1.0, 0.0, 0.0, MODE_NONE);
  };
  KPM_TRY(reverse(ysrc, Ls));

  // to the host, scaled by 1/NR, in each operator's CSC order
  std::vector<double> m(2 * (nH + nA + nB) * chunks);
  KPM_HIP_TRY(hipMemcpyAsync(m.data(), MH, gbytes, hipMemcpyDeviceToHost, ctx->stream));
  ++ctx->counters[KPM_COUNT_SYNC];
  KPM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const double inv = 1.0 / (double)NR;
  auto out = [&](const kpm_csr* A, size_t off, size_t n, double* g) {
    for (int64_t k = 0; k < A->nnz; ++k) {
      const size_t p = off + (size_t)A->csc_to_csr[k];
      double gr = 0.0, gi = 0.0;
      for (int c = 0; c < chunks; ++c) {
        gr += m[2 * (p + c * n)];
        gi += m[2 * (p + c * n) + 1];
      }
      g[2 * k] = gr * inv;
      g[2 * k + 1] = gi * inv;
    }
  };
  out(H, 0, nH, grad_H);
  out(Ja, nH * chunks, nA, grad_Ja);
  out(Jb, (nH + nA) * chunks, nB, grad_Jb);
  return KPM_OK;
}

}  // extern "C"
