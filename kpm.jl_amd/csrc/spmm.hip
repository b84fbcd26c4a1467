// spmm.hip -- the fused Chebyshev recurrence pass for gfx950.
//
// One launch is one recurrence step of KPM.jl's chebyshev_iter_single
// (src/utils/chebyshev_iteration.jl:59-75; CUSPARSE SpMM in
// ext/KPMCUDAExt.jl:84-88) over all probe columns at once, with the moment
// dot products of columnwise_dot! (src/utils/vectors.jl:103-117;
// ext/KPMCUDAExt.jl:128-145) reduced in the same pass:
//
//   out[i,:] = alpha * sum_k A[i,k] x[k,:] + beta * z[i,:]
//   MODE_DBL: d0[:] += conj(x[i,:]) x[i,:],  d1[:] += conj(x[i,:]) out[i,:]
//   MODE_LR : d0[:] += conj(l[i,:]) out[i,:]
//
// Mapping (wave64):
//   A row is NRp complex values = RW 16-byte words.  LPR lanes cover one row
//   (lane sub-index = word), so a wave covers 64/LPR rows at once and every
//   neighbour gather is one contiguous LPR*16-byte load.  Each lane keeps
//   UNR rows in flight and issues up to 4 gathers per row per batch.
//   The lane <-> probe-column mapping is fixed, so the dot partial sums need
//   no cross-lane shuffles inside the row loop; they are reduced once per
//   workgroup through LDS in a fixed order and written as per-workgroup
//   partials, which launch_combine reduces in a fixed order (deterministic:
//   same inputs -> bitwise same moments).
//   Workgroups are persistent and XCD-aware: MI355X dispatches workgroup b to
//   XCD b % 8, so XCD k owns a contiguous eighth of the row tiles and its
//   workgroups sweep that range together, keeping the neighbour rows of the
//   sweep front in that XCD's L2.
//   Arithmetic is FP64 for both storage precisions; FP32 blocks are rounded
//   on store and the dots use the stored (rounded) values.
#include "kpm_internal.h"

#include <algorithm>

namespace kpm {

template <int PREC>
struct Word;

template <>
struct Word<KPM_VEC_F64> {
  using T = double2;
  static constexpr int CPL = 1;
  __device__ static inline void unpack(const T& w, double* re, double* im) {
    re[0] = w.x;
    im[0] = w.y;
  }
  __device__ static inline T pack(const double* re, const double* im) {
    return make_double2(re[0], im[0]);
  }
  __device__ static inline T zero() { return make_double2(0.0, 0.0); }
};

template <>
struct Word<KPM_VEC_F32> {
  using T = float4;
  static constexpr int CPL = 2;
  __device__ static inline void unpack(const T& w, double* re, double* im) {
    re[0] = (double)w.x;
    im[0] = (double)w.y;
    re[1] = (double)w.z;
    im[1] = (double)w.w;
  }
  __device__ static inline T pack(const double* re, const double* im) {
    return make_float4((float)re[0], (float)im[0], (float)re[1], (float)im[1]);
  }
  __device__ static inline T zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
};

constexpr int kBlock = 256;  // 4 waves

// Per-lane metadata of one row tile: the lane's rows, their first nonzero and
// nonzero count, and the column indices of the first KB nonzeros.
template <int UNR>
struct RowMeta {
  int row[UNR], s[UNR], cnt[UNR];
};
template <int UNR, int KB>
struct ColMeta {
  int c[UNR][KB];
};

// Software pipeline over the tiles of a workgroup: while the gathers of
// tile t are in flight the lane already loads the column indices of tile t+1
// and the row pointers of tile t+2, so each tile costs one memory latency
// instead of three dependent ones (row pointer -> column index -> gather).
// Nonzeros beyond the first KB of a row are handled by a plain tail loop.
template <int PREC, int LPR, int UNR, int KB, bool REALH, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_step(const int* __restrict__ rowptr, const int* __restrict__ colind,
           const void* __restrict__ vals_, const typename Word<PREC>::T* __restrict__ x,
           const typename Word<PREC>::T* z, typename Word<PREC>::T* out,
           const typename Word<PREC>::T* __restrict__ l, double2* __restrict__ partials, int NH,
           int RW, int NRp, int ntiles, double alpha, double beta) {
  using WT = Word<PREC>;
  using T = typename WT::T;
  constexpr int CPL = WT::CPL;
  constexpr int RPW = 64 / LPR;
  constexpr int TR = 4 * RPW * UNR;
  constexpr int ND = MODE == MODE_DBL ? 2 : (MODE == MODE_LR ? 1 : 0);
  constexpr int NDA = ND > 0 ? ND : 1;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int sub = lane % LPR;
  const int rslot = lane / LPR;
  const int wcol = blockIdx.y * LPR + sub;
  const int rbase = wave * (RPW * UNR) + rslot;

  const double2* cvals = reinterpret_cast<const double2*>(vals_);
  const double* rvals = reinterpret_cast<const double*>(vals_);

  double accr[NDA][CPL], acci[NDA][CPL];
#pragma unroll
  for (int d = 0; d < NDA; ++d)
#pragma unroll
    for (int c = 0; c < CPL; ++c) accr[d][c] = acci[d][c] = 0.0;

  const int G = gridDim.x;
  const int b = blockIdx.x;
  const int xcd = b & 7;
  const int q = b >> 3;
  const int Bx = G >> 3;
  const int t0 = (int)(((int64_t)ntiles * xcd) >> 3);
  const int t1 = (int)(((int64_t)ntiles * (xcd + 1)) >> 3);

  auto load_rows = [&](int t, RowMeta<UNR>& R) {
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int row = t * TR + rbase + u * RPW;
      R.row[u] = row;
      if (t < t1 && row < NH) {
        const int s0 = rowptr[row];
        R.s[u] = s0;
        R.cnt[u] = rowptr[row + 1] - s0;
      } else {
        R.s[u] = 0;
        R.cnt[u] = 0;
      }
    }
  };
  auto load_cols = [&](const RowMeta<UNR>& R, ColMeta<UNR, KB>& Cm) {
#pragma unroll
    for (int u = 0; u < UNR; ++u)
#pragma unroll
      for (int jj = 0; jj < KB; ++jj) Cm.c[u][jj] = jj < R.cnt[u] ? colind[R.s[u] + jj] : 0;
  };
  auto mac = [&](double hr, double hi, const T& xw, double* yr, double* yi) {
    double xr[CPL], xi[CPL];
    WT::unpack(xw, xr, xi);
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      if constexpr (REALH) {
        yr[c] = fma(hr, xr[c], yr[c]);
        yi[c] = fma(hr, xi[c], yi[c]);
      } else {
        yr[c] = fma(hr, xr[c], fma(-hi, xi[c], yr[c]));
        yi[c] = fma(hr, xi[c], fma(hi, xr[c], yi[c]));
      }
    }
  };

  int t = t0 + q;
  RowMeta<UNR> R0, R1;
  ColMeta<UNR, KB> C0;
  load_rows(t, R0);
  load_cols(R0, C0);
  load_rows(t + Bx, R1);

  for (; t < t1; t += Bx) {
    // 1. this tile's gathers, values and row-local operands
    T xv[UNR][KB];
    double hr[UNR][KB], hi[UNR][KB];
#pragma unroll
    for (int u = 0; u < UNR; ++u)
#pragma unroll
      for (int jj = 0; jj < KB; ++jj) {
        if (jj < R0.cnt[u]) {
          const int k = R0.s[u] + jj;
          if constexpr (REALH) {
            hr[u][jj] = rvals[k];
            hi[u][jj] = 0.0;
          } else {
            const double2 h = cvals[k];
            hr[u][jj] = h.x;
            hi[u][jj] = h.y;
          }
          xv[u][jj] = x[(size_t)C0.c[u][jj] * RW + wcol];
        } else {
          hr[u][jj] = 0.0;
          hi[u][jj] = 0.0;
          xv[u][jj] = WT::zero();
        }
      }
    T zv[UNR], sv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const bool live = R0.row[u] < NH && t < t1;
      const size_t off = (size_t)(live ? R0.row[u] : 0) * RW + wcol;
      zv[u] = (beta != 0.0 && live) ? z[off] : WT::zero();
      if constexpr (MODE == MODE_DBL) sv[u] = live ? x[off] : WT::zero();
      if constexpr (MODE == MODE_LR) sv[u] = live ? l[off] : WT::zero();
    }
    // 2. prefetch: column indices of the next tile, row pointers of the one after
    ColMeta<UNR, KB> C1;
    load_cols(R1, C1);
    RowMeta<UNR> R2;
    load_rows(t + 2 * Bx, R2);

    // 3. accumulate
    double yr[UNR][CPL], yi[UNR][CPL];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
#pragma unroll
      for (int c = 0; c < CPL; ++c) yr[u][c] = yi[u][c] = 0.0;
#pragma unroll
      for (int jj = 0; jj < KB; ++jj) mac(hr[u][jj], hi[u][jj], xv[u][jj], yr[u], yi[u]);
      for (int j = KB; j < R0.cnt[u]; ++j) {  // rows longer than KB
        const int k = R0.s[u] + j;
        double h_r, h_i;
        if constexpr (REALH) {
          h_r = rvals[k];
          h_i = 0.0;
        } else {
          const double2 h = cvals[k];
          h_r = h.x;
          h_i = h.y;
        }
        mac(h_r, h_i, x[(size_t)colind[k] * RW + wcol], yr[u], yi[u]);
      }
    }

    // 4. combine with z, store, fused dots
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      if (R0.row[u] >= NH) continue;
      const size_t off = (size_t)R0.row[u] * RW + wcol;
      double orr[CPL], oii[CPL];
      if (beta != 0.0) {
        double zr[CPL], zi[CPL];
        WT::unpack(zv[u], zr, zi);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          orr[c] = fma(alpha, yr[u][c], beta * zr[c]);
          oii[c] = fma(alpha, yi[u][c], beta * zi[c]);
        }
      } else {
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          orr[c] = alpha * yr[u][c];
          oii[c] = alpha * yi[u][c];
        }
      }
      const T ow = WT::pack(orr, oii);
      out[off] = ow;
      if constexpr (MODE != MODE_NONE) {
        double sr[CPL], si[CPL], ar[CPL], ai[CPL];
        WT::unpack(ow, sr, si);
        WT::unpack(sv[u], ar, ai);
        if constexpr (MODE == MODE_DBL) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            accr[0][c] = fma(ar[c], ar[c], fma(ai[c], ai[c], accr[0][c]));
            accr[1][c] = fma(ar[c], sr[c], fma(ai[c], si[c], accr[1][c]));
            acci[1][c] = fma(ar[c], si[c], fma(-ai[c], sr[c], acci[1][c]));
          }
        } else {
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            accr[0][c] = fma(ar[c], sr[c], fma(ai[c], si[c], accr[0][c]));
            acci[0][c] = fma(ar[c], si[c], fma(-ai[c], sr[c], acci[0][c]));
          }
        }
      }
    }
    R0 = R1;
    C0 = C1;
    R1 = R2;
  }

  if constexpr (ND > 0) {
    __shared__ double2 red[kBlock * NDA * CPL];
#pragma unroll
    for (int d = 0; d < NDA; ++d)
#pragma unroll
      for (int c = 0; c < CPL; ++c) red[(d * CPL + c) * kBlock + tid] = make_double2(accr[d][c], acci[d][c]);
    __syncthreads();
    constexpr int nout = LPR * CPL * ND;
    for (int o = tid; o < nout; o += kBlock) {
      const int d = o / (LPR * CPL);
      const int rem = o % (LPR * CPL);
      const int sb = rem / CPL;
      const int c = rem % CPL;
      double sr = 0.0, si = 0.0;
      for (int w = 0; w < 4; ++w)
        for (int rs = 0; rs < RPW; ++rs) {
          const double2 v = red[(d * CPL + c) * kBlock + w * 64 + rs * LPR + sb];
          sr += v.x;
          si += v.y;
        }
      const int col = (blockIdx.y * LPR + sb) * CPL + c;
      partials[((size_t)b * ND + d) * NRp + col] = make_double2(sr, si);
    }
  }
}

// d0 = <a|b> per column (the bare columnwise_dot! reduction)
template <int PREC, int LPR>
__global__ void __launch_bounds__(kBlock)
    k_dot(const typename Word<PREC>::T* __restrict__ a, const typename Word<PREC>::T* __restrict__ bv,
          double2* __restrict__ partials, int NH, int RW, int NRp, int ntiles) {
  using WT = Word<PREC>;
  constexpr int CPL = WT::CPL;
  constexpr int RPW = 64 / LPR;
  constexpr int TR = 4 * RPW;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int sub = lane % LPR;
  const int rslot = lane / LPR;
  const int wcol = blockIdx.y * LPR + sub;
  double accr[CPL], acci[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) accr[c] = acci[c] = 0.0;
  const int G = gridDim.x;
  const int b = blockIdx.x;
  const int xcd = b & 7, q = b >> 3, Bx = G >> 3;
  const int t0 = (int)(((int64_t)ntiles * xcd) >> 3);
  const int t1 = (int)(((int64_t)ntiles * (xcd + 1)) >> 3);
  for (int t = t0 + q; t < t1; t += Bx) {
    const int row = t * TR + wave * RPW + rslot;
    if (row >= NH) continue;
    const size_t off = (size_t)row * RW + wcol;
    double ar[CPL], ai[CPL], br[CPL], bi[CPL];
    WT::unpack(a[off], ar, ai);
    WT::unpack(bv[off], br, bi);
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      accr[c] = fma(ar[c], br[c], fma(ai[c], bi[c], accr[c]));
      acci[c] = fma(ar[c], bi[c], fma(-ai[c], br[c], acci[c]));
    }
  }
  __shared__ double2 red[kBlock * CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) red[c * kBlock + tid] = make_double2(accr[c], acci[c]);
  __syncthreads();
  constexpr int nout = LPR * CPL;
  for (int o = tid; o < nout; o += kBlock) {
    const int sb = o / CPL;
    const int c = o % CPL;
    double sr = 0.0, si = 0.0;
    for (int w = 0; w < 4; ++w)
      for (int rs = 0; rs < RPW; ++rs) {
        const double2 v = red[c * kBlock + w * 64 + rs * LPR + sb];
        sr += v.x;
        si += v.y;
      }
    const int col = (blockIdx.y * LPR + sb) * CPL + c;
    partials[(size_t)b * NRp + col] = make_double2(sr, si);
  }
}

__global__ void __launch_bounds__(kBlock)
    k_combine(const double2* __restrict__ partials, int G, int nd, int NRp, int NR,
              double2* __restrict__ mu, Combine cb) {
  const int d = blockIdx.x / NR;
  const int col = blockIdx.x % NR;
  double sr = 0.0, si = 0.0;
  for (int g = threadIdx.x; g < G; g += kBlock) {
    const double2 v = partials[((size_t)g * nd + d) * NRp + col];
    sr += v.x;
    si += v.y;
  }
  __shared__ double2 sm[kBlock];
  sm[threadIdx.x] = make_double2(sr, si);
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double2 o = sm[threadIdx.x + w];
      sm[threadIdx.x].x += o.x;
      sm[threadIdx.x].y += o.y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double2 tot = sm[0];
    double rr = cb.alpha[d] * tot.x;
    double ri = cb.alpha[d] * tot.y;
    if (cb.bcol[d] >= 0) {
      const double2 bv = mu[col + (size_t)NR * cb.bcol[d]];
      rr = fma(cb.bscale[d], bv.x, rr);
      ri = fma(cb.bscale[d], bv.y, ri);
    }
    mu[col + (size_t)NR * cb.tcol[d]] = make_double2(rr, ri);
  }
}

// host column-major ComplexF64 (NH x NR) -> row-interleaved block (NH x NRp)
template <int PREC>
__global__ void k_upload(const double2* __restrict__ src, void* __restrict__ dst_, int64_t NH,
                         int64_t NR, int64_t NRp) {
  const int64_t total = NH * NRp;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = idx / NRp;
    const int64_t col = idx - row * NRp;
    const double2 v = col < NR ? src[col * NH + row] : make_double2(0.0, 0.0);
    if constexpr (PREC == KPM_VEC_F64) {
      reinterpret_cast<double2*>(dst_)[idx] = v;
    } else {
      reinterpret_cast<float2*>(dst_)[idx] = make_float2((float)v.x, (float)v.y);
    }
  }
}

template <int PREC>
__global__ void k_download(const void* __restrict__ src_, double2* __restrict__ dst, int64_t NH,
                           int64_t NR, int64_t NRp) {
  const int64_t total = NH * NR;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t col = idx / NH;
    const int64_t row = idx - col * NH;
    double2 v;
    if constexpr (PREC == KPM_VEC_F64) {
      v = reinterpret_cast<const double2*>(src_)[row * NRp + col];
    } else {
      const float2 f = reinterpret_cast<const float2*>(src_)[row * NRp + col];
      v = make_double2((double)f.x, (double)f.y);
    }
    dst[idx] = v;
  }
}

// plain streaming copy of n 16-byte words: the calibration kernel for the
// HBM byte counters and the practical bandwidth ceiling of the device
__global__ void __launch_bounds__(kBlock) k_copy16(const float4* __restrict__ src,
                                                   float4* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    dst[i] = src[i];
}

int launch_copy16(kpm_ctx* ctx, const void* src, void* dst, int64_t nwords) {
  const int64_t g = std::min<int64_t>((nwords + kBlock - 1) / kBlock,
                                      (int64_t)std::max(ctx->num_cus, 8) * 16);
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)std::max<int64_t>(g, 1)), dim3(kBlock), 0,
                     ctx->stream, (const float4*)src, (float4*)dst, nwords);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

// ---------------------------------------------------------------------------
// dispatch tables

using StepFn = void (*)(const int*, const int*, const void*, const void*, const void*, void*,
                        const void*, double2*, int, int, int, int, double, double);

template <int PREC, int LPR, int UNR, int KB, bool REALH, int MODE>
static void launch_step_t(dim3 grid, hipStream_t st, const int* rp, const int* ci, const void* v,
                          const void* x, const void* z, void* out, const void* l, double2* part,
                          int NH, int RW, int NRp, int ntiles, double alpha, double beta) {
  using T = typename Word<PREC>::T;
  hipLaunchKernelGGL((k_step<PREC, LPR, UNR, KB, REALH, MODE>), grid, dim3(kBlock), 0, st, rp, ci, v,
                     (const T*)x, (const T*)z, (T*)out, (const T*)l, part, NH, RW, NRp, ntiles,
                     alpha, beta);
}

using StepLauncher = void (*)(dim3, hipStream_t, const int*, const int*, const void*, const void*,
                              const void*, void*, const void*, double2*, int, int, int, int,
                              double, double);

template <int PREC, int LPR, int UNR, int KB, bool REALH>
static StepLauncher pick_mode(int mode) {
  switch (mode) {
    case MODE_DBL: return &launch_step_t<PREC, LPR, UNR, KB, REALH, MODE_DBL>;
    case MODE_LR: return &launch_step_t<PREC, LPR, UNR, KB, REALH, MODE_LR>;
    default: return &launch_step_t<PREC, LPR, UNR, KB, REALH, MODE_NONE>;
  }
}

template <int PREC, int LPR, int UNR>
static StepLauncher pick_real(int kb, bool realh, int mode) {
  if (kb <= 4)
    return realh ? pick_mode<PREC, LPR, UNR, 4, true>(mode) : pick_mode<PREC, LPR, UNR, 4, false>(mode);
  return realh ? pick_mode<PREC, LPR, UNR, 8, true>(mode) : pick_mode<PREC, LPR, UNR, 8, false>(mode);
}

template <int PREC, int LPR>
static StepLauncher pick_unr(int unr, int kb, bool realh, int mode) {
  return unr == 1 ? pick_real<PREC, LPR, 1>(kb, realh, mode) : pick_real<PREC, LPR, 2>(kb, realh, mode);
}

template <int PREC>
static StepLauncher pick_lpr(int lpr, int unr, int kb, bool realh, int mode) {
  switch (lpr) {
    case 1: return pick_unr<PREC, 1>(unr, kb, realh, mode);
    case 2: return pick_unr<PREC, 2>(unr, kb, realh, mode);
    case 4: return pick_unr<PREC, 4>(unr, kb, realh, mode);
    case 8: return pick_unr<PREC, 8>(unr, kb, realh, mode);
    case 16: return pick_unr<PREC, 16>(unr, kb, realh, mode);
    case 32: return pick_unr<PREC, 32>(unr, kb, realh, mode);
    default: return pick_unr<PREC, 64>(unr, kb, realh, mode);
  }
}

int row_words(int64_t NR, int prec) {
  const int64_t cpw = cplx_per_word(prec);
  int64_t w = (NR + cpw - 1) / cpw;
  if (w <= 64) {
    int64_t p = 1;
    while (p < w) p <<= 1;
    return (int)p;
  }
  return (int)((w + 63) / 64 * 64);
}

static int default_grid(kpm_ctx* ctx, int64_t ntiles) {
  const int64_t cap = (int64_t)std::max(ctx->num_cus, 8) * 8;  // 8 workgroups of 4 waves per CU
  int64_t g = std::min<int64_t>(ntiles, cap);
  g = (g + 7) / 8 * 8;
  return (int)std::max<int64_t>(g, 8);
}

int launch_step(kpm_ctx* ctx, const StepLaunch& s, const StepConfig& cfg, int* grid_used) {
  const kpm_block* x = s.x;
  const int prec = x->prec;
  const int RW = row_words(x->NR, prec);
  const int lpr = std::min(RW, 64);
  const int chunks = RW / lpr;
  const int unr = cfg.unroll == 1 ? 1 : 2;
  const int kb = s.A->max_row_nnz <= 4 ? 4 : 8;
  const int rpw = 64 / lpr;
  const int64_t TR = 4LL * rpw * unr;
  const int64_t ntiles = (x->NH + TR - 1) / TR;
  if (ntiles > INT32_MAX) return fail(KPM_EINVAL, "row count too large for the recurrence kernel");
  int G = cfg.grid > 0 ? (cfg.grid + 7) / 8 * 8 : default_grid(ctx, ntiles);
  const int nd = s.mode == MODE_DBL ? 2 : (s.mode == MODE_LR ? 1 : 0);
  if (nd > 0) KPM_TRY(ensure(ctx, ctx->partials, (size_t)G * nd * x->NRp * sizeof(double2)));
  StepLauncher fn = prec == KPM_VEC_F32 ? pick_lpr<KPM_VEC_F32>(lpr, unr, kb, s.A->is_real != 0, s.mode)
                                        : pick_lpr<KPM_VEC_F64>(lpr, unr, kb, s.A->is_real != 0, s.mode);
  fn(dim3(G, chunks), ctx->stream, s.A->rowptr, s.A->colind, s.A->vals, x->d, s.z ? s.z->d : nullptr,
     s.out->d, s.l ? s.l->d : nullptr, (double2*)ctx->partials.ptr, (int)x->NH, RW, (int)x->NRp,
     (int)ntiles, s.alpha, s.beta);
  KPM_HIP_TRY(hipGetLastError());
  if (grid_used) *grid_used = G;
  return KPM_OK;
}

template <int PREC, int LPR>
static void launch_dot_t(dim3 grid, hipStream_t st, const void* a, const void* b, double2* part,
                         int NH, int RW, int NRp, int ntiles) {
  using T = typename Word<PREC>::T;
  hipLaunchKernelGGL((k_dot<PREC, LPR>), grid, dim3(kBlock), 0, st, (const T*)a, (const T*)b, part,
                     NH, RW, NRp, ntiles);
}

using DotLauncher = void (*)(dim3, hipStream_t, const void*, const void*, double2*, int, int, int, int);

template <int PREC>
static DotLauncher pick_dot(int lpr) {
  switch (lpr) {
    case 1: return &launch_dot_t<PREC, 1>;
    case 2: return &launch_dot_t<PREC, 2>;
    case 4: return &launch_dot_t<PREC, 4>;
    case 8: return &launch_dot_t<PREC, 8>;
    case 16: return &launch_dot_t<PREC, 16>;
    case 32: return &launch_dot_t<PREC, 32>;
    default: return &launch_dot_t<PREC, 64>;
  }
}

int launch_dot(kpm_ctx* ctx, const kpm_block* a, const kpm_block* b, int* grid_used) {
  const int prec = a->prec;
  const int RW = row_words(a->NR, prec);
  const int lpr = std::min(RW, 64);
  const int chunks = RW / lpr;
  const int64_t TR = 4LL * (64 / lpr);
  const int64_t ntiles = (a->NH + TR - 1) / TR;
  const int G = default_grid(ctx, ntiles);
  KPM_TRY(ensure(ctx, ctx->partials, (size_t)G * a->NRp * sizeof(double2)));
  DotLauncher fn = prec == KPM_VEC_F32 ? pick_dot<KPM_VEC_F32>(lpr) : pick_dot<KPM_VEC_F64>(lpr);
  fn(dim3(G, chunks), ctx->stream, a->d, b->d, (double2*)ctx->partials.ptr, (int)a->NH, RW,
     (int)a->NRp, (int)ntiles);
  KPM_HIP_TRY(hipGetLastError());
  if (grid_used) *grid_used = G;
  return KPM_OK;
}

int launch_combine(kpm_ctx* ctx, int grid, int64_t NR, int64_t NRp, const Combine& c, void* mu_dev,
                   int64_t /*NCmu*/) {
  hipLaunchKernelGGL(k_combine, dim3((unsigned)(c.nd * NR)), dim3(kBlock), 0, ctx->stream,
                     (const double2*)ctx->partials.ptr, grid, c.nd, (int)NRp, (int)NR,
                     (double2*)mu_dev, c);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

static unsigned copy_grid(int64_t total) {
  int64_t g = (total + kBlock - 1) / kBlock;
  return (unsigned)std::min<int64_t>(std::max<int64_t>(g, 1), 65536);
}

int launch_upload(kpm_ctx* ctx, const void* staging, kpm_block* b) {
  const unsigned g = copy_grid(b->NH * b->NRp);
  if (b->prec == KPM_VEC_F32)
    hipLaunchKernelGGL(k_upload<KPM_VEC_F32>, dim3(g), dim3(kBlock), 0, ctx->stream,
                       (const double2*)staging, b->d, b->NH, b->NR, b->NRp);
  else
    hipLaunchKernelGGL(k_upload<KPM_VEC_F64>, dim3(g), dim3(kBlock), 0, ctx->stream,
                       (const double2*)staging, b->d, b->NH, b->NR, b->NRp);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

int launch_download(kpm_ctx* ctx, const kpm_block* b, void* staging) {
  const unsigned g = copy_grid(b->NH * b->NR);
  if (b->prec == KPM_VEC_F32)
    hipLaunchKernelGGL(k_download<KPM_VEC_F32>, dim3(g), dim3(kBlock), 0, ctx->stream, b->d,
                       (double2*)staging, b->NH, b->NR, b->NRp);
  else
    hipLaunchKernelGGL(k_download<KPM_VEC_F64>, dim3(g), dim3(kBlock), 0, ctx->stream, b->d,
                       (double2*)staging, b->NH, b->NR, b->NRp);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

}  // namespace kpm
