// step_kernel.h -- the fused Chebyshev recurrence pass for gfx950 (kernel
// template; instantiated per storage precision in step_f64.hip / step_f32.hip).
//
// One launch is one recurrence step of KPM.jl's chebyshev_iter_single
// (src/utils/chebyshev_iteration.jl:59-75; CUSPARSE SpMM in
// ext/KPMCUDAExt.jl:84-88) over all probe columns at once, with the moment
// dot products of columnwise_dot! (src/utils/vectors.jl:103-117;
// ext/KPMCUDAExt.jl:128-145) reduced in the same pass:
//
//   out[i,:] = alpha * sum_k A[i,k] x[k,:] + beta * z[i,:]
//   MODE_DBL: d0[:] += conj(x[i,:]) x[i,:],  d1[:] += conj(x[i,:]) out[i,:]
//   MODE_DBLA: the MODE_DBL dots summed over the probe columns too (the
//             probe-averaged moments), one partial pair per wave
//   MODE_LR : d0[:] += conj(l[i,:]) out[i,:]
//   MODE_SRC: out[i,:] += gamma * l[i,:], no dots (the adjoint's source term)
// z may alias out (in place); when beta == 0 the launcher passes x as z and
// the kernel selects (never multiplies) it away.
//
// Mapping (wave64):
//   A row is NRp complex values = RW 16-byte words.  LPR lanes cover one row
//   (lane sub-index = word), so a wave covers 64/LPR rows at once and every
//   neighbour gather is one contiguous LPR*16-byte load.  Each lane keeps
//   UNR rows and KB nonzeros per row in flight.
//   The lane <-> probe-column mapping is fixed, so the dot partial sums need
//   no cross-lane shuffles inside the row loop; they are reduced once per
//   workgroup through LDS in a fixed order and written as per-workgroup
//   partials, which launch_combine reduces in a fixed order (deterministic:
//   same inputs -> bitwise same moments).
//   Workgroups are persistent and XCD-aware: MI355X dispatches workgroup b to
//   XCD b % 8, so XCD k owns a contiguous eighth of the row tiles and its
//   workgroups sweep that range together, keeping the neighbour rows of the
//   sweep front in that XCD's L2.
//   Memory phase is branch-free: every load is issued unconditionally with a
//   clamped, always-valid index and masked by a zero weight, so the compiler
//   keeps all of a tile's loads in flight together (an exec-masked branch
//   around a load forces a full vmcnt wait in front of it).
//   Software pipeline: while the gathers of tile t are in flight the lane
//   already loads the column indices of tile t+1 and the row pointers of
//   tile t+2, so a tile costs one memory latency instead of three dependent
//   ones (row pointer -> column index -> gather).
//   Arithmetic follows the storage precision: the sparse product and the
//   combination with z run in FP64 for FP64 blocks and in FP32 for FP32
//   blocks (operator values rounded to FP32 as they are loaded); the moment
//   dots always convert the stored values to FP64 and accumulate in FP64.
#pragma once

#include <type_traits>

#include "bf16_planes.h"
#include "kpm_internal.h"

namespace kpm {

template <int PREC>
struct Word;

// A: the arithmetic type of the recurrence (the sparse product and the
// combination with z).  The dots always take the stored values to FP64.
template <>
struct Word<KPM_VEC_F64> {
  using T = double2;
  using A = double;
  static constexpr int CPL = 1;
  __device__ static inline void unpack_a(const T& w, A* re, A* im) {
    re[0] = w.x;
    im[0] = w.y;
  }
  __device__ static inline T pack_a(const A* re, const A* im) { return make_double2(re[0], im[0]); }
  __device__ static inline void unpack(const T& w, double* re, double* im) {
    re[0] = w.x;
    im[0] = w.y;
  }
  __device__ static inline T pack(const double* re, const double* im) {
    return make_double2(re[0], im[0]);
  }
  __device__ static inline T zero() { return make_double2(0.0, 0.0); }
  // p ? a : b per component (a select of the whole vector type is lowered
  // through private memory)
  __device__ static inline T pick(bool p, const T& a, const T& b) {
    return make_double2(p ? a.x : b.x, p ? a.y : b.y);
  }
};

template <>
struct Word<KPM_VEC_F32> {
  using T = float4;
  using A = float;
  static constexpr int CPL = 2;
  __device__ static inline void unpack_a(const T& w, A* re, A* im) {
    re[0] = w.x;
    im[0] = w.y;
    re[1] = w.z;
    im[1] = w.w;
  }
  __device__ static inline T pack_a(const A* re, const A* im) {
    return make_float4(re[0], im[0], re[1], im[1]);
  }
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
  __device__ static inline T pick(bool p, const T& a, const T& b) {
    return make_float4(p ? a.x : b.x, p ? a.y : b.y, p ? a.z : b.z, p ? a.w : b.w);
  }
};

constexpr int kBlock = 64 * kStepWaves;  // 4 waves

// whether the fused pass reads the operator's FP32 values (kpm_csr::vals32)
__host__ __device__ constexpr bool vals32_step(int prec, int lpr, int mode) {
  return prec == KPM_VEC_F32 && lpr <= 8 && mode != MODE_DUAL && mode != MODE_DUALP;
}

// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Lane J of every aligned group of G lanes, broadcast to its group, with DPP
// moves (VALU only, no LDS traffic).  G = 16: row_newbcast (gfx90a and
// later); G = 8: row_newbcast of lane J and of lane 8 + J, selected by the
// half-row; G = 4: quad_perm [J, J, J, J]; G = 2: quad_perm
// [J, J, 2 + J, 2 + J].  Every lane of the wave must be active.
template <int G, int J>
__device__ __forceinline__ int group_bcast(int v) {
  static_assert(G == 2 || G == 4 || G == 8 || G == 16, "group of 2, 4, 8 or 16 lanes");
  if constexpr (G == 2) {
    return __builtin_amdgcn_update_dpp(0, v, J | (J << 2) | ((2 + J) << 4) | ((2 + J) << 6), 0xF,
                                       0xF, false);
  } else if constexpr (G == 4) {
    return __builtin_amdgcn_update_dpp(0, v, J * 0x55, 0xF, 0xF, false);
  } else if constexpr (G == 8) {
    const int a = __builtin_amdgcn_update_dpp(0, v, 0x150 + J, 0xF, 0xF, false);
    const int b = __builtin_amdgcn_update_dpp(0, v, 0x158 + J, 0xF, 0xF, false);
    return (threadIdx.x & 8) ? b : a;
  } else {
    return __builtin_amdgcn_update_dpp(0, v, 0x150 + J, 0xF, 0xF, false);
  }
}
// Sum over each aligned group of L lanes (L = 1 .. 64), the same value in
// every lane of the group: DPP row_mirror, row_half_mirror and the quad
// exchanges [2,3,0,1], [1,0,3,2] within 16 lanes (each step adds a lane's
// value to its partner's, so partners hold bitwise equal sums), then lane
// exchanges for 32 and 64.  Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int L>
__device__ __forceinline__ double group_sum(double v) {
  static_assert(L == 1 || L == 2 || L == 4 || L == 8 || L == 16 || L == 32 || L == 64, "group");
  if constexpr (L >= 16) v += dpp_move<0x140>(v);  // lane i + lane 15 - i of the row
  if constexpr (L >= 8) v += dpp_move<0x141>(v);   // lane i + lane 7 - i of the half-row
  if constexpr (L >= 4) v += dpp_move<0x4E>(v);    // quad [2, 3, 0, 1]
  if constexpr (L >= 2) v += dpp_move<0xB1>(v);    // quad [1, 0, 3, 2]
  if constexpr (L >= 32) v += __shfl_xor(v, 16);
  if constexpr (L >= 64) v += __shfl_xor(v, 32);
  return v;
}

template <int G, int J>
__device__ __forceinline__ float group_bcast(float v) {
  return __int_as_float(group_bcast<G, J>(__float_as_int(v)));
}
template <int G, int J>
__device__ __forceinline__ double group_bcast(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = group_bcast<G, J>((int)(b & 0xffffffffLL));
  const int hi = group_bcast<G, J>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Streamed operands (z = v_{n-1}, read once; out = v_{n+1}, not reread until
// the next pass) use non-temporal loads/stores so they do not evict the
// neighbour rows of x that the gathers reuse.  KPM_STREAM_NT=0 builds plain
// accesses instead.
#ifndef KPM_STREAM_NT
#define KPM_STREAM_NT 1
#endif
typedef double nt_d2 __attribute__((ext_vector_type(2)));
typedef float nt_f4 __attribute__((ext_vector_type(4)));

__device__ inline double2 ld_stream(const double2* p) {
#if KPM_STREAM_NT
  const nt_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_d2*>(p));
  return make_double2(v.x, v.y);
#else
  return *p;
#endif
}
__device__ inline float4 ld_stream(const float4* p) {
#if KPM_STREAM_NT
  const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
#else
  return *p;
#endif
}
__device__ inline void st_stream(double2* p, const double2& v) {
#if KPM_STREAM_NT
  __builtin_nontemporal_store(nt_d2{v.x, v.y}, reinterpret_cast<nt_d2*>(p));
#else
  *p = v;
#endif
}
__device__ inline void st_stream(float4* p, const float4& v) {
#if KPM_STREAM_NT
  __builtin_nontemporal_store(nt_f4{v.x, v.y, v.z, v.w}, reinterpret_cast<nt_f4*>(p));
#else
  *p = v;
#endif
}

// Fixed-order reduction of the moments of one dot d and a group of kCombCols
// consecutive probe columns over G per-workgroup partials stored [g][d][col],
// then columnwise_dot!'s alpha * dot + beta_scale * target[:, beta_col]
// (src/utils/vectors.jl:103-117).  Thread t takes column c0 + t % kCombCols
// and the partials of workgroups t / kCombCols, + kCombRows, ... in that
// order (loads issued 16 at a time), so the kCombCols columns of one
// workgroup's row, one 128-byte line, are read together by one unit; a
// fixed tree over the kCombRows lanes of each column finishes the sum.
// sm: kBlock double2 of LDS; every thread of the workgroup must call it.
constexpr int kCombCols = 8;
__host__ __device__ inline int combine_units(int nd, int NR) {
  return nd * ((NR + kCombCols - 1) / kCombCols);
}
// Columns one unit reduces: kCombCols, or 1 for a one-column table (the
// probe sums of MODE_DBLA), so that every thread of the unit adds rows.
__device__ inline int comb_cols(int NR) { return NR == 1 ? 1 : kCombCols; }
// Sum over partial rows [g0, g1) of one column (row stride gs): thread t
// takes column t % ncols and rows g0 + t / ncols, + kBlock / ncols, ... in
// that order (loads issued 4 at a time), then a fixed tree over the threads
// of each column.  The column's sum lands in sm[t] of its thread with
// t / ncols == 0.  sm: kBlock double2 of LDS; every thread must call it.
__device__ inline void sum_rows(const double2* __restrict__ p, int g0, int g1, size_t gs, int ncols,
                                double2* sm) {
  const int lane = (int)threadIdx.x / ncols;
  const int rows = kBlock / ncols;
  double sr = 0.0, si = 0.0;
  // loads in flight per thread: kept small because the fused pass carries
  // this code, and its register allocation is the maximum over the kernel
  // (16 in flight took the FP32 graphene pass from 107 to 148 VGPRs, 4 to 3
  // waves per SIMD)
  constexpr int kBatch = 4;
  int g = g0 + lane;
  for (; g + (kBatch - 1) * rows < g1; g += kBatch * rows) {
    double2 v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) v[u] = p[(size_t)(g + u * rows) * gs];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      sr += v[u].x;
      si += v[u].y;
    }
  }
  for (; g < g1; g += rows) {
    const double2 v = p[(size_t)g * gs];
    sr += v.x;
    si += v.y;
  }
  sm[threadIdx.x] = make_double2(sr, si);
  __syncthreads();
  for (int w = rows / 2; w > 0; w >>= 1) {  // lanes of a column are ncols apart
    if (lane < w) {
      const double2 o = sm[threadIdx.x + w * ncols];
      sm[threadIdx.x].x += o.x;
      sm[threadIdx.x].y += o.y;
    }
    __syncthreads();
  }
}

// Fixed-order reduction of the moments of one dot d and a group of
// comb_cols(NR) consecutive probe columns over G partial rows stored
// [g][d][col], then columnwise_dot!'s alpha * dot + beta_scale *
// target[:, beta_col] (src/utils/vectors.jl:103-117).  The kCombCols columns
// of one row (one 128-byte line) are read together by one unit.
// sm: kBlock double2 of LDS; every thread of the workgroup must call it.
__device__ inline void combine_unit(const double2* __restrict__ partials, int G, int nd, int NRp,
                                    int NR, double2* __restrict__ mu, const Combine& cb, int unit,
                                    double2* sm) {
  const int ncols = comb_cols(NR);
  const int ngrp = (NR + kCombCols - 1) / kCombCols;
  const int d = unit / ngrp;
  const int col = (unit % ngrp) * kCombCols + (int)threadIdx.x % ncols;
  const bool live = col < NR;
  const size_t gs = (size_t)nd * NRp;  // between consecutive partial rows
  sum_rows(partials + (size_t)d * NRp + (live ? col : 0), 0, G, gs, ncols, sm);
  if ((int)threadIdx.x < ncols && live) {
    const double2 tot = sm[threadIdx.x];
    double rr = cb.alpha[d] * tot.x;
    double ri = cb.alpha[d] * tot.y;
    if (cb.bcol[d] >= 0) {
      const double2 bv = mu[col + (size_t)NR * cb.bcol[d]];
      rr = fma(cb.bscale[d], bv.x, rr);
      ri = fma(cb.bscale[d], bv.y, ri);
    }
    mu[col + (size_t)NR * cb.tcol[d]] = make_double2(rr, ri);
  }
  __syncthreads();
}

// First level of the two-level reduction: one unit sums, for one dot d and
// a group of columns, the partials of the kSlice rows of one slice and writes
// the slice sum to out[slice][d][col].
__host__ __device__ inline int reduce1_units(int nd, int NR, int G) {
  return combine_units(nd, NR) * ((G + kSlice - 1) / kSlice);
}
__device__ inline void reduce1_unit(const double2* __restrict__ partials, int G, int nd, int NRp,
                                    int NR, double2* __restrict__ out, int unit, double2* sm) {
  const int ncols = comb_cols(NR);
  const int nu = combine_units(nd, NR);
  const int slice = unit / nu;
  const int cu = unit % nu;
  const int ngrp = (NR + kCombCols - 1) / kCombCols;
  const int d = cu / ngrp;
  const int col = (cu % ngrp) * kCombCols + (int)threadIdx.x % ncols;
  const bool live = col < NR;
  const size_t gs = (size_t)nd * NRp;
  sum_rows(partials + (size_t)d * NRp + (live ? col : 0), slice * kSlice, min(G, (slice + 1) * kSlice),
           gs, ncols, sm);
  if ((int)threadIdx.x < ncols && live) out[(size_t)slice * gs + (size_t)d * NRp + col] = sm[threadIdx.x];
  __syncthreads();
}

template <int UNR>
struct RowMeta {
  int row[UNR], s[UNR], cnt[UNR];
};
template <int UNR, int KB>
struct ColMeta {
  int c[UNR][KB];
};
// Nonzeros spread over the lanes of a row: with G = min(LPR, 16) lanes per
// row inside a DPP row, lane j holds nonzeros jE .. jE + E - 1 of its row,
// column index and value, one contiguous run per array (vector loads instead
// of KB same-address loads per lane; on 2-lane rows of 7 entries, E = 4 words
// in one load instead of four); group_bcast hands slot jj to every lane of
// the row.  Rows wider than 16 lanes load the same nonzeros in each DPP row.
// DUAL adds the values of a second operator on the same pattern.
template <int UNR, int E, bool DUAL = false, class AT = double>
struct LaneMeta {
  int c[UNR][E];
  AT hr[UNR][E], hi[UNR][E];
  AT gr[DUAL ? UNR : 1][DUAL ? E : 1], gi[DUAL ? UNR : 1][DUAL ? E : 1];
};

// Segmented schedule (SEG = true, batched realisations): workgroup b owns the
// contiguous tiles [tile_lo[b], tile_lo[b+1]); treal[t] is the realisation of
// tile t (tiles never straddle realisations), and the workgroup flushes its
// dot partials to slot seg0[b] + k whenever the realisation changes.  The
// change is uniform over the workgroup, so the barrier inside the flush is
// reached by all of its threads.
//
// Dual pass (MODE == MODE_DUAL): vals2_ holds the complex values of a second
// operator B on the same pattern, and the pass also writes out2 = gamma * B x
// from the gathers of x it already holds.  With A and B merged onto the union
// of their patterns, each output sums its own nonzeros in its own row order
// (the other operator's entries weigh exactly zero), so out and out2 are those
// of two separate passes.  o2 places out2's words: word w of a row goes to
// block w / o2.cw, o2.qs words apart, at row * o2.rw + w % o2.cw (one block
// of RW words per row is the layout of out; narrower blocks split a wide row
// into separate arrays).
//
// flags: bit 0 (kStepTail), rows longer than KB exist (the tail loop runs);
// bit 1 (kStepDiagGather), every row of the operator stores its diagonal among
// its first KB nonzeros, so the doubling modes take the own-row words x[row]
// that their dots need from the gather of that slot instead of loading them
// again (the slot whose column equals the row; the load it saves is a sixth of
// the pass's vector loads on 4-entry rows); bit 2 (kStepPlanes, MODE_NONE),
// out2 also receives the split-BF16 planes (bf16_planes.h) of out, and bit 3
// (kStepPlanesOnly) leaves out unwritten.  MODE_DUALP is MODE_DUAL with the
// second product written as planes instead of its words (a mode of its own:
// a runtime branch there raised the dual pass from 68 to 86 VGPRs).
template <int PREC, int LPR, int UNR, int KB, bool REALH, int MODE, bool SEG>
__global__ void __launch_bounds__(kBlock)
    k_step(const int* __restrict__ rowptr, const int* __restrict__ colind,
           const void* __restrict__ vals_, const typename Word<PREC>::T* __restrict__ x,
           const typename Word<PREC>::T* z, typename Word<PREC>::T* out,
           const typename Word<PREC>::T* __restrict__ l, double2* __restrict__ partials, int NH,
           int RW, int NRp, int ntiles, double alpha, double beta, int flags,
           const int* __restrict__ tile_lo, const int* __restrict__ treal,
           const int* __restrict__ seg0, double gamma, PendingCombine pc,
           const void* __restrict__ vals2_, typename Word<PREC>::T* __restrict__ out2,
           SplitOut o2, int chunk, int xmap) {
  using WT = Word<PREC>;
  using T = typename WT::T;
  constexpr int CPL = WT::CPL;
  constexpr int RPW = 64 / LPR;
  constexpr int TR = 4 * RPW * UNR;
  constexpr bool kAvg = MODE == MODE_DBLA;
  constexpr bool kDbl = MODE == MODE_DBL || kAvg;
  constexpr int ND = kDbl ? 2 : (MODE == MODE_LR ? 1 : 0);
  constexpr int NDA = ND > 0 ? ND : 1;
  // rows of 2 or more lanes take their nonzeros spread over the lanes
  constexpr int MG = LPR >= 16 ? 16 : LPR;
  constexpr bool kLaneMeta = MG >= 2 && KB <= 16;
  constexpr int ME = kLaneMeta ? (KB + MG - 1) / MG : 1;
  constexpr bool kDual = MODE == MODE_DUAL || MODE == MODE_DUALP;
  constexpr bool kDualPlanes = MODE == MODE_DUALP;  // the second product as split-BF16 planes
  static_assert(!kDual || (kLaneMeta && !REALH && !SEG), "dual pass: complex values, spread metadata");
  using AT = typename WT::A;
  using CMeta = std::conditional_t<kLaneMeta, LaneMeta<UNR, ME, kDual, AT>, ColMeta<UNR, KB>>;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int sub = lane % LPR;
  const int rslot = lane / LPR;
  const int wcol = blockIdx.y * LPR + sub;
  const int rbase = wave * (RPW * UNR) + rslot;
  const bool has_z = beta != 0.0;  // uniform
  const bool has_tail = (flags & kStepTail) != 0;
  const bool dgat = kDbl && (flags & kStepDiagGather) != 0;
  // split-BF16 planes (bf16_planes.h) at out2: of out (MODE_NONE; without out
  // when planes_only) or of the second product (MODE_DUALP)
  const bool planes = MODE == MODE_NONE && !SEG && (flags & kStepPlanes) != 0;
  const bool planes_only = planes && (flags & kStepPlanesOnly) != 0;
  auto store_planes = [&](uint32_t* P, int row, const AT* re, const AT* im) {
#pragma unroll
    for (int c = 0; c < CPL; ++c)
      store_plane_value(P, (int64_t)row * NRp + (int64_t)wcol * CPL + c, (double)re[c], (double)im[c]);
  };
  size_t o2col = 0;
  if constexpr (kDual) {
    const int qb = wcol / o2.cw;
    o2col = (size_t)qb * o2.qs + (size_t)(wcol - qb * o2.cw);
  }

  // operator values: FP32 copies for FP32 blocks of rows up to 8 words (the
  // recurrence multiplies in FP32 there, so this is the same rounding as
  // converting on load, at half the value bytes), FP64 otherwise and for the
  // dual pass's merged operator.  Narrow rows are where the values are a
  // large share of the pass's bytes (cubic NR = 4: 84 of 184 bytes per row
  // with FP64 values); on 32-word rows (graphene NR = 64) the FP32 copies
  // measured 0.5% slower, so those keep the FP64 values.  vals32_step()
  // makes the launchers' choice.
  constexpr bool kVal32 = vals32_step(PREC, LPR, MODE);
  using RV = std::conditional_t<kVal32, float, double>;
  using CV = std::conditional_t<kVal32, float2, double2>;
  const CV* cvals = reinterpret_cast<const CV*>(vals_);
  const RV* rvals = reinterpret_cast<const RV*>(vals_);
  const double2* cvals2 = reinterpret_cast<const double2*>(vals2_);

  double accr[NDA][CPL], acci[NDA][CPL];
#pragma unroll
  for (int d = 0; d < NDA; ++d)
#pragma unroll
    for (int c = 0; c < CPL; ++c) accr[d][c] = acci[d][c] = 0.0;

  const int G = gridDim.x;
  const int b = blockIdx.x;
  int tstart, t1, Bx;
  if constexpr (SEG) {
    tstart = tile_lo[b];
    t1 = tile_lo[b + 1];
    Bx = 1;
  } else if (chunk > 0) {
    // one-shot schedule: workgroup b sweeps chunk c of `chunk` consecutive
    // tiles.  xmap = 0: c = b, so the dispatcher's in-order hand-out keeps
    // the whole grid on one front; xmap = 1: XCD k (= b % 8) takes the
    // chunks of its own contiguous eighth in order (one front per XCD, the
    // neighbour rows of a front in that XCD's L2); xmap = m >= 2: one front,
    // but XCD k takes runs of m consecutive chunks (round r of 8m
    // workgroups covers chunks [8mr, 8m(r+1)), XCD k the k-th run), so a
    // row's near neighbours sit in the same XCD's L2 except at every m-th
    // chunk boundary (a short last round uses shorter runs; G is a multiple
    // of 8).  Chunks past the last tile are empty and only write zero
    // partials.
    int c = b;
    if (xmap == 1) {
      c = (b & 7) * (G >> 3) + (b >> 3);
    } else if (xmap > 1) {
      const int r0 = b / (8 * xmap) * (8 * xmap);
      const int q = min(xmap, (G - r0) >> 3);
      c = r0 + (b & 7) * q + ((b - r0) >> 3);
    }
    tstart = (int)min<int64_t>((int64_t)c * chunk, ntiles);
    t1 = (int)min<int64_t>((int64_t)tstart + chunk, ntiles);
    Bx = 1;
  } else {
    const int xcd = b & 7;
    Bx = G >> 3;
    tstart = (int)(((int64_t)ntiles * xcd) >> 3) + (b >> 3);
    t1 = (int)(((int64_t)ntiles * (xcd + 1)) >> 3);
  }

  // row pointers of tile t (rows beyond the matrix or the range read row 0
  // and get count 0).  Rows of 16 or 32 lanes: the wave's RPW * UNR
  // consecutive rows take their RPW * UNR + 1 row pointers with wave-uniform
  // (scalar) loads, off the vector-memory path, and each row group selects
  // its pair.  (Measured: 5% on 16-lane rows, 2% on 32-lane rows, and 4%
  // slower on 64-lane rows, which keep the vector loads.)
  auto load_rows = [&](int t, RowMeta<UNR>& R) {
    if constexpr (LPR == 16 || LPR == 32) {
      constexpr int NP = RPW * UNR + 1;
      const int r0 = __builtin_amdgcn_readfirstlane(t * TR + wave * (RPW * UNR));
      int pv[NP];
#pragma unroll
      for (int k = 0; k < NP; ++k) pv[k] = rowptr[min(r0 + k, NH)];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int idx = rslot + u * RPW;
        int s0 = pv[0], e0 = pv[1];
#pragma unroll
        for (int k = 1; k < NP - 1; ++k) {
          s0 = idx == k ? pv[k] : s0;
          e0 = idx == k ? pv[k + 1] : e0;
        }
        const int row = r0 + idx;
        const bool v = (t < t1) & (row < NH);
        R.row[u] = row;
        R.s[u] = v ? s0 : 0;
        R.cnt[u] = v ? e0 - s0 : 0;
      }
      return;
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int row = t * TR + rbase + u * RPW;
      const bool v = (t < t1) & (row < NH);
      const int rc = v ? row : 0;
      const int s0 = rowptr[rc];
      const int e0 = rowptr[rc + 1];
      R.row[u] = row;
      R.s[u] = s0;
      R.cnt[u] = v ? e0 - s0 : 0;
    }
  };
  // column indices of the first KB nonzeros (invalid slots: index 0); with
  // kLaneMeta also their values (invalid slots: weight 0)
  auto load_cols = [&](const RowMeta<UNR>& R, CMeta& Cm) {
    if constexpr (kLaneMeta) {
      // lane j of the row holds its nonzeros j*ME .. j*ME + ME - 1, read as
      // one run per array (the compiler merges the consecutive loads into
      // vector loads; the arrays' zero tail, kCsrPad entries, covers the
      // reads past a row's end, and rows past the range read from entry 0)
      const int j = lane & (MG - 1);
      const int i0 = j * ME;
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int k0 = R.s[u] + i0;
        int cw[ME];
#pragma unroll
        for (int e = 0; e < ME; ++e) cw[e] = colind[k0 + e];
#pragma unroll
        for (int e = 0; e < ME; ++e) {
          const bool v = i0 + e < R.cnt[u];
          Cm.c[u][e] = v ? cw[e] : 0;
          if constexpr (REALH) {
            const RV h = rvals[k0 + e];
            Cm.hr[u][e] = v ? (AT)h : (AT)0;
            Cm.hi[u][e] = (AT)0;
          } else {
            const CV h = cvals[k0 + e];
            Cm.hr[u][e] = v ? (AT)h.x : (AT)0;
            Cm.hi[u][e] = v ? (AT)h.y : (AT)0;
          }
          if constexpr (kDual) {
            const double2 g = cvals2[k0 + e];
            Cm.gr[u][e] = v ? (AT)g.x : (AT)0;
            Cm.gi[u][e] = v ? (AT)g.y : (AT)0;
          }
        }
      }
    } else {
      // one lane per row: the row's first KB column indices as one run (the
      // zero tail covers the reads past its end)
#pragma unroll
      for (int u = 0; u < UNR; ++u)
#pragma unroll
        for (int jj = 0; jj < KB; ++jj) {
          const int c = colind[R.s[u] + jj];
          Cm.c[u][jj] = jj < R.cnt[u] ? c : 0;
        }
    }
  };
  auto mac = [&](AT hr, AT hi, const T& xw, AT* yr, AT* yi) {
    AT xr[CPL], xi[CPL];
    WT::unpack_a(xw, xr, xi);
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

  // fixed-order workgroup reduction of the per-lane dot sums into dst
  // (MODE_DBLA uses it only as the pending reductions' kBlock-entry scratch)
  __shared__ double2 red[ND > 0 ? (kAvg ? kBlock : kBlock * NDA * CPL) : 1];
  // (dot d, column col) goes to dst[(d * NRp + col) * stride]
  auto flush = [&](double2* dst, size_t stride) {
    if constexpr (kAvg) return;  // per-wave partials instead (below)
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
      dst[((size_t)d * NRp + col) * stride] = make_double2(sr, si);
    }
    __syncthreads();
  };
  // the previous step's moment columns (uniform per workgroup)
  if constexpr (!SEG && ND > 0) {
    // the previous steps' reductions, uniform per workgroup: final units
    // first, then first-level units
    const int nfin = pc.part ? combine_units(pc.cb.nd, pc.NR) : 0;
    const int nl1 = pc.l1_part ? reduce1_units(pc.l1_nd, pc.NR, pc.l1_G) : 0;
    for (int q = blockIdx.y * gridDim.x + blockIdx.x; q < nfin + nl1; q += gridDim.x * gridDim.y) {
      if (q < nfin)
        combine_unit(pc.part, pc.G, pc.cb.nd, pc.NRp, pc.NR, pc.mu, pc.cb, q, red);
      else
        reduce1_unit(pc.l1_part, pc.l1_G, pc.l1_nd, pc.NRp, pc.NR, pc.l1_out, q - nfin, red);
    }
  }
  int seg = 0, cur = 0;
  if constexpr (SEG) {
    seg = seg0[b];
    cur = tstart < t1 ? treal[tstart] : -1;
  }

  int t = tstart;
  RowMeta<UNR> R0, R1;
  CMeta C0;
  load_rows(t, R0);
  load_cols(R0, C0);
  load_rows(t + Bx, R1);

  for (; t < t1; t += Bx) {
    // 1. this tile's values, gathers and row-local operands, all in flight
    AT hr[UNR][KB], hi[UNR][KB];
    AT gr[kDual ? UNR : 1][kDual ? KB : 1], gi[kDual ? UNR : 1][kDual ? KB : 1];
    T xv[UNR][KB];
    bool own[kDbl ? UNR : 1][kDbl ? KB : 1];  // slot jj of row u is the row's own column
    if constexpr (kLaneMeta) {  // every lane is active here (uniform loop)
#pragma unroll
      for (int u = 0; u < UNR; ++u)
        static_for<0, KB>([&](auto J) {
          constexpr int jj = decltype(J)::value;
          constexpr int e = jj % ME, src = jj / ME;  // nonzero jj: lane jj / ME, entry jj % ME
          const int c = group_bcast<MG, src>(C0.c[u][e]);
          if constexpr (kDbl) own[u][jj] = c == R0.row[u];
          hr[u][jj] = group_bcast<MG, src>(C0.hr[u][e]);
          hi[u][jj] = REALH ? (AT)0 : group_bcast<MG, src>(C0.hi[u][e]);
          if constexpr (kDual) {
            gr[u][jj] = group_bcast<MG, src>(C0.gr[u][e]);
            gi[u][jj] = group_bcast<MG, src>(C0.gi[u][e]);
          }
          xv[u][jj] = x[(size_t)c * RW + wcol];
        });
    } else {
#pragma unroll
      for (int u = 0; u < UNR; ++u)
#pragma unroll
        for (int jj = 0; jj < KB; ++jj) {
          const bool v = jj < R0.cnt[u];
          const int k = R0.s[u] + jj;  // one run per row (zero tail past its end)
          if constexpr (REALH) {
            const RV h = rvals[k];
            hr[u][jj] = v ? (AT)h : (AT)0;
            hi[u][jj] = (AT)0;
          } else {
            const CV h = cvals[k];
            hr[u][jj] = v ? (AT)h.x : (AT)0;
            hi[u][jj] = v ? (AT)h.y : (AT)0;
          }
          xv[u][jj] = x[(size_t)C0.c[u][jj] * RW + wcol];
          if constexpr (kDbl) own[u][jj] = C0.c[u][jj] == R0.row[u];
        }
    }
    T zv[UNR], sv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int rc = R0.row[u] < NH ? R0.row[u] : 0;
      const size_t off = (size_t)rc * RW + wcol;
      zv[u] = ld_stream(z + off);  // z is always a valid block; unused when beta == 0
      if constexpr (kDbl) {
        if (!dgat) sv[u] = x[off];
      }
      if constexpr (MODE == MODE_LR || MODE == MODE_SRC) sv[u] = l[off];
    }
    // 2. prefetch: column indices of the next tile, row pointers of the one after
    CMeta C1;
    load_cols(R1, C1);
    RowMeta<UNR> R2;
    load_rows(t + 2 * Bx, R2);

    // 3. accumulate
    AT yr[UNR][CPL], yi[UNR][CPL];
    AT y2r[kDual ? UNR : 1][CPL], y2i[kDual ? UNR : 1][CPL];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
#pragma unroll
      for (int c = 0; c < CPL; ++c) yr[u][c] = yi[u][c] = (AT)0;
#pragma unroll
      for (int jj = 0; jj < KB; ++jj) mac(hr[u][jj], hi[u][jj], xv[u][jj], yr[u], yi[u]);
      if constexpr (kDual) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) y2r[u][c] = y2i[u][c] = (AT)0;
#pragma unroll
        for (int jj = 0; jj < KB; ++jj) mac(gr[u][jj], gi[u][jj], xv[u][jj], y2r[u], y2i[u]);
      }
    }
    if (has_tail) {  // uniform: rows longer than KB exist in this operator
#pragma unroll
      for (int u = 0; u < UNR; ++u)
        for (int j = KB; j < R0.cnt[u]; ++j) {
          const int k = R0.s[u] + j;
          AT h_r, h_i;
          if constexpr (REALH) {
            h_r = (AT)rvals[k];
            h_i = (AT)0;
          } else {
            const CV h = cvals[k];
            h_r = (AT)h.x;
            h_i = (AT)h.y;
          }
          const T xw = x[(size_t)colind[k] * RW + wcol];
          mac(h_r, h_i, xw, yr[u], yi[u]);
          if constexpr (kDual) {
            const double2 g = cvals2[k];
            mac((AT)g.x, (AT)g.y, xw, y2r[u], y2i[u]);
          }
        }
    }

    if constexpr (kDbl) {
      if (dgat) {  // own-row words from the gathered slot (rows past NH: any slot, unused)
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
          sv[u] = xv[u][0];
#pragma unroll
          for (int jj = 1; jj < KB; ++jj) sv[u] = WT::pick(own[u][jj], xv[u][jj], sv[u]);
        }
      }
    }

    // 4. combine with z, store, fused dots
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const bool live = R0.row[u] < NH;
      const AT al = (AT)alpha, be = (AT)beta, ga = (AT)gamma;
      AT orr[CPL], oii[CPL], zr[CPL], zi[CPL];
      WT::unpack_a(zv[u], zr, zi);
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        orr[c] = has_z ? fma(al, yr[u][c], be * zr[c]) : al * yr[u][c];
        oii[c] = has_z ? fma(al, yi[u][c], be * zi[c]) : al * yi[u][c];
      }
      if constexpr (MODE == MODE_SRC) {  // + gamma * s (reverse Clenshaw source)
        AT qr[CPL], qi[CPL];
        WT::unpack_a(sv[u], qr, qi);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          orr[c] = fma(ga, qr[c], orr[c]);
          oii[c] = fma(ga, qi[c], oii[c]);
        }
      }
      const T ow = WT::pack_a(orr, oii);
      if (live && !planes_only) st_stream(out + (size_t)R0.row[u] * RW + wcol, ow);
      if constexpr (MODE == MODE_NONE && !SEG) {
        if (planes && live) store_planes(reinterpret_cast<uint32_t*>(out2), R0.row[u], orr, oii);
      }
      if constexpr (kDual) {
        AT pr[CPL], pi[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          pr[c] = ga * y2r[u][c];
          pi[c] = ga * y2i[u][c];
        }
        if (live) {
          if constexpr (kDualPlanes) store_planes(reinterpret_cast<uint32_t*>(out2), R0.row[u], pr, pi);
          else st_stream(out2 + (size_t)R0.row[u] * o2.rw + o2col, WT::pack_a(pr, pi));
        }
      }
      if constexpr (kDbl || MODE == MODE_LR) {
        double sr[CPL], si[CPL], ar[CPL], ai[CPL];
        WT::unpack(ow, sr, si);
        WT::unpack(sv[u], ar, ai);
        const double w = live ? 1.0 : 0.0;
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          ar[c] *= w;
          ai[c] *= w;
        }
        if constexpr (kDbl) {
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
    if constexpr (SEG && ND > 0) {
      const int nxt = t + 1 < t1 ? treal[t + 1] : -1;  // uniform over the workgroup
      if (nxt != cur) {
        flush(partials + (size_t)seg * ND * NRp, 1);
#pragma unroll
        for (int d = 0; d < NDA; ++d)
#pragma unroll
          for (int c = 0; c < CPL; ++c) accr[d][c] = acci[d][c] = 0.0;
        ++seg;
        cur = nxt;
      }
    }
  }

  if constexpr (kAvg) {
    // probe-summed dots: the wave's rows and columns summed with DPP moves in
    // a fixed order, one partial pair per wave at [wave id][d] (no LDS
    // round trip and no barrier; 32 bytes per wave instead of 32 * NRp per
    // workgroup)
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      s0 += accr[0][c];
      s1 += accr[1][c];
      s2 += acci[1][c];
    }
    s0 = group_sum<64>(s0);
    s1 = group_sum<64>(s1);
    s2 = group_sum<64>(s2);
    const size_t wid = ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (kBlock / 64) + wave;
    if (lane == 0) {
      partials[wid * 2] = make_double2(s0, 0.0);
      partials[wid * 2 + 1] = make_double2(s1, s2);
    }
  } else if constexpr (!SEG && ND > 0) {
    // per-workgroup partials [b][d][col]: each workgroup writes whole lines
    flush(partials + (size_t)b * ND * NRp, 1);
  }
}

// ---- launcher tables -------------------------------------------------------------

using StepLauncher = void (*)(dim3, hipStream_t, const int*, const int*, const void*, const void*,
                              const void*, void*, const void*, double2*, int, int, int, int,
                              double, double, int, const int*, const int*, const int*, double,
                              const PendingCombine&, const void*, void*, const SplitOut&, int, int);

template <int PREC, int LPR, int UNR, int KB, bool REALH, int MODE, bool SEG>
static void launch_step_t(dim3 grid, hipStream_t st, const int* rp, const int* ci, const void* v,
                          const void* x, const void* z, void* out, const void* l, double2* part,
                          int NH, int RW, int NRp, int ntiles, double alpha, double beta,
                          int has_tail, const int* tile_lo, const int* treal, const int* seg0,
                          double gamma, const PendingCombine& pc, const void* v2, void* out2,
                          const SplitOut& o2, int chunk, int xmap) {
  using T = typename Word<PREC>::T;
  hipLaunchKernelGGL((k_step<PREC, LPR, UNR, KB, REALH, MODE, SEG>), grid, dim3(kBlock), 0, st, rp,
                     ci, v, (const T*)x, (const T*)z, (T*)out, (const T*)l, part, NH, RW, NRp,
                     ntiles, alpha, beta, has_tail, tile_lo, treal, seg0, gamma, pc, v2, (T*)out2, o2,
                     chunk, xmap);
}

template <int PREC, int LPR, int UNR, int KB, bool REALH>
static StepLauncher pick_mode(int mode) {
  switch (mode) {
    case MODE_DBL: return &launch_step_t<PREC, LPR, UNR, KB, REALH, MODE_DBL, false>;
    case MODE_DBLA: return &launch_step_t<PREC, LPR, UNR, KB, REALH, MODE_DBLA, false>;
    case MODE_LR: return &launch_step_t<PREC, LPR, UNR, KB, REALH, MODE_LR, false>;
    case MODE_DBL_SEG: return &launch_step_t<PREC, LPR, UNR, KB, REALH, MODE_DBL, true>;
    case MODE_SRC: return &launch_step_t<PREC, LPR, UNR, KB, REALH, MODE_SRC, false>;
    case MODE_DUAL:  // complex merged values on the spread-metadata path only
      if constexpr (!REALH && LPR >= 2) return &launch_step_t<PREC, LPR, UNR, KB, false, MODE_DUAL, false>;
      return nullptr;
    case MODE_DUALP:
      if constexpr (!REALH && LPR >= 2) return &launch_step_t<PREC, LPR, UNR, KB, false, MODE_DUALP, false>;
      return nullptr;
    default: return &launch_step_t<PREC, LPR, UNR, KB, REALH, MODE_NONE, false>;
  }
}

template <int PREC, int LPR, int UNR, int KB>
static StepLauncher pick_real(bool realh, int mode) {
  return realh ? pick_mode<PREC, LPR, UNR, KB, true>(mode) : pick_mode<PREC, LPR, UNR, KB, false>(mode);
}

// Gather slots per row: 4, 5, 6, 7, 8, 10, 12 or 16, the smallest that holds
// the operator's longest row (slots past a row's end still issue a gather of
// row 0, so the common lattices get their own widths: the cubic lattice's 7
// entries per row at width 7 instead of 8, batched ensemble of 256
// realisations 1.597e12 -> 1.642e12 moments/s; the disordered square
// lattice's 5 at width 5 instead of 6, DOS-loss gradient on 512^2 133 -> 130
// ms).  KB 10, 12 and 16 keep one row in flight per lane (register budget).
template <int PREC, int LPR>
static StepLauncher pick_kb(int unr, int kb, bool realh, int mode) {
  if (kb <= 4)
    return unr == 1 ? pick_real<PREC, LPR, 1, 4>(realh, mode) : pick_real<PREC, LPR, 2, 4>(realh, mode);
  if (kb <= 5)
    return unr == 1 ? pick_real<PREC, LPR, 1, 5>(realh, mode) : pick_real<PREC, LPR, 2, 5>(realh, mode);
  if (kb <= 6)
    return unr == 1 ? pick_real<PREC, LPR, 1, 6>(realh, mode) : pick_real<PREC, LPR, 2, 6>(realh, mode);
  if (kb <= 7)
    return unr == 1 ? pick_real<PREC, LPR, 1, 7>(realh, mode) : pick_real<PREC, LPR, 2, 7>(realh, mode);
  if (kb <= 8)
    return unr == 1 ? pick_real<PREC, LPR, 1, 8>(realh, mode) : pick_real<PREC, LPR, 2, 8>(realh, mode);
  if (kb <= 10) return pick_real<PREC, LPR, 1, 10>(realh, mode);
  if (kb <= 12) return pick_real<PREC, LPR, 1, 12>(realh, mode);
  return pick_real<PREC, LPR, 1, 16>(realh, mode);
}

template <int PREC>
static StepLauncher pick_step(int lpr, int unr, int kb, bool realh, int mode) {
  switch (lpr) {
    case 1: return pick_kb<PREC, 1>(unr, kb, realh, mode);
    case 2: return pick_kb<PREC, 2>(unr, kb, realh, mode);
    case 4: return pick_kb<PREC, 4>(unr, kb, realh, mode);
    case 8: return pick_kb<PREC, 8>(unr, kb, realh, mode);
    case 16: return pick_kb<PREC, 16>(unr, kb, realh, mode);
    case 32: return pick_kb<PREC, 32>(unr, kb, realh, mode);
    default: return pick_kb<PREC, 64>(unr, kb, realh, mode);
  }
}

// defined in step_f64.hip / step_f32.hip
StepLauncher pick_step_f64(int lpr, int unr, int kb, bool realh, int mode);
StepLauncher pick_step_f32(int lpr, int unr, int kb, bool realh, int mode);

}  // namespace kpm
