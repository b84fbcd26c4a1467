// spmm.hip -- launchers of the fused recurrence pass (kernel template in
// step_kernel.h), the dot-only pass, the fixed-order moment reduction, the
// layout conversions between host column-major blocks and device
// row-interleaved blocks, and the calibration copy kernel.
#include "step_kernel.h"

#include <algorithm>

namespace kpm {

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
  __shared__ double2 sm[kBlock];
  combine_unit(partials, G, nd, NRp, NR, mu, cb, blockIdx.x, sm);
}

__global__ void __launch_bounds__(kBlock)
    k_reduce1(const double2* __restrict__ partials, int G, int nd, int NRp, int NR,
              double2* __restrict__ out) {
  __shared__ double2 sm[kBlock];
  reduce1_unit(partials, G, nd, NRp, NR, out, blockIdx.x, sm);
}

int launch_reduce1(kpm_ctx* ctx, const double2* partials, int G, int nd, int64_t NR, int64_t NRp,
                   double2* out) {
  hipLaunchKernelGGL(k_reduce1, dim3((unsigned)reduce1_units(nd, (int)NR, G)), dim3(kBlock), 0,
                     ctx->stream, partials, G, nd, (int)NRp, (int)NR, out);
  KPM_HIP_TRY(hipGetLastError());
  ++ctx->counters[KPM_COUNT_COMBINE];
  return KPM_OK;
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
// HBM byte counters and the practical bandwidth ceiling of the device.  One
// word per lane over a full grid: the dispatcher hands out workgroups in
// address order, so the in-flight window stays one contiguous front (6.3 TB/s
// on MI355X for 1 GiB operands, against 4.7-5.5 TB/s for a persistent
// grid-stride loop, whose workgroups drift apart).
__global__ void __launch_bounds__(kBlock) k_copy16(const float4* __restrict__ src,
                                                   float4* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    dst[i] = src[i];
}

int launch_copy16(kpm_ctx* ctx, const void* src, void* dst, int64_t nwords) {
  const int64_t g = std::min<int64_t>((nwords + kBlock - 1) / kBlock, (int64_t)INT32_MAX);
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)std::max<int64_t>(g, 1)), dim3(kBlock), 0,
                     ctx->stream, (const float4*)src, (float4*)dst, nwords);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

// ---------------------------------------------------------------------------
// dispatch tables

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

int default_grid(kpm_ctx* ctx, int64_t ntiles) {
  const int64_t cap = (int64_t)std::max(ctx->num_cus, 8) * 8;  // 8 workgroups of 4 waves per CU
  int64_t g = std::min<int64_t>(ntiles, cap);
  g = (g + 7) / 8 * 8;
  return (int)std::max<int64_t>(g, 8);
}

StepGeom step_geom(const kpm_csr* A, int64_t NR, int prec, const StepConfig& cfg) {
  StepGeom g;
  g.RW = row_words(NR, prec);
  g.lpr = std::min(g.RW, 64);
  g.chunks = g.RW / g.lpr;
  const int w = A->max_row_nnz;
  g.kb = w <= 4 ? 4 : w <= 5 ? 5 : w <= 6 ? 6 : w <= 7 ? 7 : w <= 8 ? 8 : w <= 10 ? 10 : w <= 12 ? 12 : 16;
  g.unr = (cfg.unroll == 1 || g.kb > 8) ? 1 : 2;
  g.has_tail = A->max_row_nnz > g.kb ? 1 : 0;
  g.TR = 4LL * (64 / g.lpr) * g.unr;
  return g;
}

// the fused pass's flag bits for operator A at geometry g (the kernel reads
// kStepDiagGather only in its doubling modes)
static int step_flags(const kpm_csr* A, const StepGeom& g) {
  return (g.has_tail ? kStepTail : 0) | (A->diag_slots <= g.kb ? kStepDiagGather : 0);
}

static StepLauncher pick(int prec, const StepGeom& g, bool realh, int mode) {
  return prec == KPM_VEC_F32 ? pick_step_f32(g.lpr, g.unr, g.kb, realh, mode)
                             : pick_step_f64(g.lpr, g.unr, g.kb, realh, mode);
}

int partials_grid(kpm_ctx* ctx, const kpm_csr* A, int64_t NH, int64_t NR, int prec, int mode) {
  return partials_grid_cfg(ctx, A, NH, NR, prec, tuned_config(ctx, A, NR, prec, mode));
}

int partials_grid_cfg(kpm_ctx* ctx, const kpm_csr* A, int64_t NH, int64_t NR, int prec,
                      const StepConfig& cfg) {
  const int64_t base = (int64_t)kMaxGridPerCU * std::max(ctx->num_cus, 1);
  if (cfg.chunk <= 0) return (int)base;
  const StepGeom g = step_geom(A, NR, prec, cfg);
  const int64_t ntiles = (NH + g.TR - 1) / g.TR;
  const int64_t gc = ((ntiles + cfg.chunk - 1) / cfg.chunk + 7) / 8 * 8;
  const int64_t lim = (int64_t)kMaxChunkGridPerCU * std::max(ctx->num_cus, 1);
  return (int)std::max(base, std::min(gc, lim));
}

int launch_step(kpm_ctx* ctx, const StepLaunch& s, const StepConfig& cfg, int* grid_used) {
  const kpm_block* x = s.x;
  const int prec = x->prec;
  const StepGeom g = step_geom(s.A, x->NR, prec, cfg);
  const int64_t ntiles = (x->NH + g.TR - 1) / g.TR;
  if (ntiles > INT32_MAX) return fail(KPM_EINVAL, "row count too large for the recurrence kernel");
  int G = cfg.grid > 0 ? (cfg.grid + 7) / 8 * 8 : default_grid(ctx, ntiles);
  const bool avg = s.mode == MODE_DBLA;
  const int nd = (s.mode == MODE_DBL || avg) ? 2 : (s.mode == MODE_LR ? 1 : 0);
  // partial rows per workgroup: one per wave and row chunk for MODE_DBLA
  // (one column each), one per workgroup (NRp columns) otherwise
  const int upw = avg ? kStepWaves * g.chunks : 1;
  const int64_t pcols = avg ? 1 : x->NRp;
  double2* part = s.part_out;
  // s.max_grid counts partial rows; in workgroups:
  const int cap = (s.max_grid > 0 ? s.max_grid : kMaxGridPerCU * std::max(ctx->num_cus, 1) * upw) / upw;
  // one-shot schedule: one workgroup per chunk of tiles (a multiple of 8
  // workgroups), unless the caller's partials buffer cannot hold them
  int chunk = 0;
  if (cfg.chunk > 0) {
    const int64_t gc = ((ntiles + cfg.chunk - 1) / cfg.chunk + 7) / 8 * 8;
    if (gc <= INT32_MAX && (!part || gc <= cap)) {
      G = (int)gc;
      chunk = cfg.chunk;
    }
  }
  if (part) {
    // caller-owned buffer sized for `cap` workgroups: a persistent grid is
    // clamped to it, a one-shot grid was only taken above if it fits
    if (chunk == 0) G = std::max(8, std::min(G, cap / 8 * 8));
  } else {
    if (nd > 0) KPM_TRY(ensure(ctx, ctx->partials, (size_t)G * upw * nd * pcols * sizeof(double2)));
    part = (double2*)ctx->partials.ptr;
  }
  if ((s.mode == MODE_LR || s.mode == MODE_SRC) && !s.l) return fail(KPM_EINVAL, "missing operand block");
  if (s.mode == MODE_DUAL && (!s.vals2 || (!s.out2 && !s.planes) || s.A->is_real))
    return fail(KPM_EINVAL, "dual pass needs complex merged values and a second output");
  SplitOut o2{g.RW, g.RW, 0};
  if (s.mode == MODE_DUAL && s.out2_cols > 0) {
    const int64_t cpw = cplx_per_word(prec);
    if (s.out2_cols % cpw != 0 || s.out2_stride % cpw != 0 || s.out2_cols > g.RW * cpw ||
        (g.RW * cpw) % s.out2_cols != 0)
      return fail(KPM_EINVAL, "dual pass: second output blocks do not tile the row");
    o2 = SplitOut{(int)(s.out2_cols / cpw), (int)(s.out2_cols / cpw), s.out2_stride / cpw};
  }
  int flags = step_flags(s.A, g);
  int kmode = s.mode;
  void* out2 = s.out2 ? s.out2->d : nullptr;
  if (s.planes) {
    if (s.mode != MODE_NONE && s.mode != MODE_DUAL) return fail(KPM_EINVAL, "planes from a doubling pass");
    if (s.mode == MODE_DUAL && s.out2_cols > 0) return fail(KPM_EINVAL, "planes of a split second output");
    if (s.mode == MODE_DUAL) kmode = MODE_DUALP;
    else flags |= kStepPlanes | (s.planes_only ? kStepPlanesOnly : 0);
    out2 = s.planes;
  }
  const PendingCombine pc = s.pend ? *s.pend : PendingCombine{};
  StepLauncher fn = pick(prec, g, s.A->is_real != 0, kmode);
  if (!fn) return fail(KPM_EINVAL, "no dual pass for rows of one lane");
  // FP32 blocks of narrow rows read the operator's FP32 values
  const void* vals = vals32_step(prec, g.lpr, kmode) ? s.A->vals32 : s.A->vals;
  if (!vals) return fail(KPM_EINVAL, "operator has no FP32 values");
  fn(dim3(G, g.chunks), ctx->stream, s.A->rowptr, s.A->colind, vals, x->d, s.z ? s.z->d : x->d,
     s.out->d, s.l ? s.l->d : nullptr, part, (int)x->NH, g.RW, (int)x->NRp,
     (int)ntiles, s.alpha, s.beta, flags, nullptr, nullptr, nullptr, s.gamma, pc, s.vals2, out2, o2,
     chunk, chunk > 0 ? cfg.xmap : 0);
  KPM_HIP_TRY(hipGetLastError());
  ++ctx->counters[KPM_COUNT_STEP];
  if (s.mode == MODE_DUAL) ++ctx->counters[KPM_COUNT_DUAL];
  if (grid_used) *grid_used = G * upw;  // partial rows written
  return KPM_OK;
}

int launch_step_seg(kpm_ctx* ctx, const StepLaunch& s, const StepGeom& g, int G, int64_t ntiles,
                    const int* tile_lo, const int* treal, const int* seg0, double2* partials) {
  const kpm_block* x = s.x;
  StepLauncher fn = pick(x->prec, g, s.A->is_real != 0, MODE_DBL_SEG);
  const void* vals = vals32_step(x->prec, g.lpr, MODE_DBL_SEG) ? s.A->vals32 : s.A->vals;
  if (!vals) return fail(KPM_EINVAL, "operator has no FP32 values");
  fn(dim3(G, g.chunks), ctx->stream, s.A->rowptr, s.A->colind, vals, x->d, s.z ? s.z->d : x->d,
     s.out->d, nullptr, partials, (int)x->NH, g.RW, (int)x->NRp, (int)ntiles, s.alpha, s.beta,
     step_flags(s.A, g), tile_lo, treal, seg0, 0.0, PendingCombine{}, nullptr, nullptr, SplitOut{}, 0,
     0);
  KPM_HIP_TRY(hipGetLastError());
  ++ctx->counters[KPM_COUNT_STEP];
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
  ++ctx->counters[KPM_COUNT_DOT];
  if (grid_used) *grid_used = G;
  return KPM_OK;
}

int launch_combine(kpm_ctx* ctx, int grid, int64_t NR, int64_t NRp, const Combine& c, void* mu_dev,
                   int64_t /*NCmu*/) {
  return launch_combine_from(ctx, (const double2*)ctx->partials.ptr, grid, NR, NRp, c, mu_dev);
}

int launch_combine_from(kpm_ctx* ctx, const double2* partials, int grid, int64_t NR, int64_t NRp,
                        const Combine& c, void* mu_dev) {
  hipLaunchKernelGGL(k_combine, dim3((unsigned)combine_units(c.nd, (int)NR)), dim3(kBlock), 0, ctx->stream, partials,
                     grid, c.nd, (int)NRp, (int)NR, (double2*)mu_dev, c);
  KPM_HIP_TRY(hipGetLastError());
  ++ctx->counters[KPM_COUNT_COMBINE];
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
