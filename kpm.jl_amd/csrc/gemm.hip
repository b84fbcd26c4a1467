// gemm.hip -- the kpm_2d moment-table contraction on the MFMA cores.
//
// Replaces the ZGEMM mu_b = R^H L / NR of _kpm_2d_blocked!
// (src/moment.jl:1055-1061) and the host accumulation mu[n,m] += conj(mu_b)
// with the parity mask (src/moment.jl:1062-1072):
//
//   mu[n0+n, m0+m] += scale * sum_k conj(L_m[k]) R_n[k]      (k over NH*NRp)
//
// Slots are contiguous recurrence blocks (row-interleaved NH x NRp complex),
// so a slot is a dense vector of K complex = 2K reals; the padded probe
// columns are zero in every slot and add nothing.  The complex product runs
// as two real GEMMs over the interleaved reals:
//   Re = sum_k' R'[n,k'] L'[m,k']            L' = (Lr, Li) interleaved
//   Im = sum_k' R'[n,k'] L''[m,k']           L'' = (-Li, Lr)
// L'' is formed in registers from the pair partner, so each operand is read
// once.  Each lane takes 8 consecutive reals of its row per 32-real stage from
// LDS and feeds them to 8 MFMAs (the k order inside a stage is a fixed
// permutation, identical for both operands).
//
// Instructions: v_mfma_f64_16x16x4f64 (FP64 mode) and v_mfma_f32_16x16x4f32
// (FP32 mode: FP32 accumulation over runs of 256 reals of K, promoted into
// FP64 accumulators after each run).  Split-K partials are
// stored in FP64 and reduced in a fixed order, then scaled, masked and added.
//
// Operand layout of a 16x16x4 MFMA (wave64): lane l supplies A[l%16][l/16]
// and B[l/16][l%16].  It receives, in accumulator element v, D[4*(l/16)+v][l%16]
// for the f32 instruction and D[l/16+4*v][l%16] for the f64 instruction
// (measured on gfx950 with an in-bounds probe).
#include "kpm_internal.h"

#include <algorithm>

#include "bf16_planes.h"

namespace kpm {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename MT>
struct Mfma;

template <>
struct Mfma<double> {
  using acc_t = d4;
  // output row of accumulator element v for lane group g = l/16
  __device__ static inline int row(int g, int v) { return g + 4 * v; }
  __device__ static inline acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
  __device__ static inline acc_t mac(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

template <>
struct Mfma<float> {
  using acc_t = f4;
  __device__ static inline int row(int g, int v) { return 4 * g + v; }
  __device__ static inline acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  __device__ static inline acc_t mac(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

constexpr int kStep = 32;  // reals of K per stage (16 complex values per slot row)

// Workgroup tile: 64 (n) x 64 (m), four waves of 32 x 32 (2 x 2 MFMA tiles).
// Each stage stages the 64 R rows and 64 L rows of one 32-real K segment in
// LDS with coalesced loads (consecutive lanes read consecutive 16-byte words
// of one row), converted to the MFMA type; waves then read their MFMA
// operands from LDS.  Rows are padded by 16 bytes so the 16 lanes of an
// MFMA lane group, which read 16 different rows at the same offset, hit 16
// different bank quads.  Two LDS stages: the global loads of stage s+1 are in
// flight while stage s is multiplied.
// VT: storage real type of the slots; MT: MFMA arithmetic type.
template <typename VT, typename MT>
__global__ void __launch_bounds__(256, 2)
    k_contract(const VT* __restrict__ L, int nL, const VT* __restrict__ R, int nR, int64_t K2,
               int64_t P2, int64_t kchunk, int tm, int ntiles, int splits,
               double2* __restrict__ part, int ldn, int ldm) {
  using MF = Mfma<MT>;
  using acc_t = typename MF::acc_t;
  constexpr int PADE = 16 / (int)sizeof(MT);       // pad elements per LDS row
  constexpr int LDR = kStep + PADE;                 // LDS row stride (MT elements)
  constexpr int WPR = kStep * (int)sizeof(VT) / 16; // 16-byte global words per row segment
  constexpr int EPW = 16 / (int)sizeof(VT);         // VT elements per 16-byte word
  constexpr int NW = 2 * 64 * WPR / 256;            // global words per thread per stage (A+B)
  __shared__ MT lds[2][2][64 * LDR];                // [stage][A|B][row * LDR + k]

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int g = lane >> 4;
  const int i = lane & 15;
  // Workgroup b runs on XCD b % 8.  With splits a multiple of 8, XCD x takes
  // the K-chunks z = x, x + 8, ... and, for each, all tiles of the table in
  // consecutive dispatch slots, so the tiles resident together on one XCD
  // read the same K-chunk of every slot and each slot segment is fetched
  // into that XCD's L2 once rather than once per tile.
  int tile, z;
  if ((splits & 7) == 0) {
    const int b = blockIdx.x, local = b >> 3;
    tile = local % ntiles;
    z = (local / ntiles) * 8 + (b & 7);
  } else {
    tile = blockIdx.x % ntiles;
    z = blockIdx.x / ntiles;
  }
  const int tile_n = (tile / tm) * 64, tile_m = (tile % tm) * 64;
  const int n0 = (wave >> 1) * 32;
  const int m0 = (wave & 1) * 32;
  const int64_t kbeg = (int64_t)z * kchunk;
  const int64_t kend = min(kbeg + kchunk, K2);
  const int nstage = kend > kbeg ? (int)((kend - kbeg) / kStep) : 0;

  // this thread's global words: word w -> (operand, row, word-in-row)
  const VT* src[NW];
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    const int w = tid + 256 * q;
    const int op = w / (64 * WPR);  // 0: R (A operand), 1: L (B operand)
    const int rem = w % (64 * WPR);
    const int row = rem / WPR;
    const int wi = rem % WPR;
    const int grow = op == 0 ? min(tile_n + row, nR - 1) : min(tile_m + row, nL - 1);
    src[q] = (op == 0 ? R : L) + (size_t)grow * P2 + kbeg + wi * EPW;
  }
  auto gload = [&](int stage, float4* regs) {
#pragma unroll
    for (int q = 0; q < NW; ++q) regs[q] = *reinterpret_cast<const float4*>(src[q] + (int64_t)stage * kStep);
  };
  auto lstore = [&](int buf, const float4* regs) {
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      const int w = tid + 256 * q;
      const int op = w / (64 * WPR);
      const int rem = w % (64 * WPR);
      const int row = rem / WPR;
      const int wi = rem % WPR;
      const VT* e = reinterpret_cast<const VT*>(&regs[q]);
      MT* dst = &lds[buf][op][row * LDR + wi * EPW];
#pragma unroll
      for (int t = 0; t < EPW; ++t) dst[t] = (MT)e[t];
    }
  };

  // FP64 MFMA: the 3M complex product (3 real products per complex MAC
  // instead of 4): are = sum Ar Lr, aim = sum Ai Li, a3 = sum (Ar+Ai)(Lr-Li);
  // Re = are + aim, Im = a3 - are + aim.  Other types: interleaved 4M form,
  // are / aim being Re / Im directly.
  constexpr bool k3M = sizeof(MT) == 8;
  acc_t are[2][2], aim[2][2], a3[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) are[a][b] = aim[a][b] = a3[a][b] = MF::zero();

  // FP32 MFMA mode: the FP32 accumulators are added into FP64 ones every
  // kFlush stages (256 reals of K), so FP32 rounding accumulates over short
  // runs only; FP64 mode accumulates in the MFMA registers directly.
  constexpr bool kPromote = sizeof(MT) == 4;
  constexpr int kFlush = 8;
  double fre[2][2][4], fim[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) fre[a][b][v] = fim[a][b][v] = 0.0;
  auto flush = [&]() {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          fre[a][b][v] += (double)are[a][b][v];
          fim[a][b][v] += (double)aim[a][b][v];
        }
        are[a][b] = aim[a][b] = MF::zero();
      }
  };

  float4 regs[NW];
  if (nstage > 0) {
    gload(0, regs);
    lstore(0, regs);
  }
  __syncthreads();
  for (int st = 0; st < nstage; ++st) {
    const int buf = st & 1;
    if (st + 1 < nstage) gload(st + 1, regs);  // next stage in flight during the MFMAs
    MT av[2][8], bv[2][8];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const MT* pa = &lds[buf][0][(n0 + 16 * t + i) * LDR + 8 * g];
      const MT* pb = &lds[buf][1][(m0 + 16 * t + i) * LDR + 8 * g];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        av[t][j] = pa[j];
        bv[t][j] = pb[j];
      }
    }
    if constexpr (k3M) {
      // per complex element c of the lane's 8 reals: P1 += Ar Lr, P2 += Ai Li,
      // P3 += (Ar + Ai)(Lr - Li)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int tm = 0; tm < 2; ++tm) {
          const MT lr = bv[tm][2 * c], li = bv[tm][2 * c + 1];
          const MT ld = lr - li;
#pragma unroll
          for (int tn = 0; tn < 2; ++tn) {
            const MT ar = av[tn][2 * c], ai = av[tn][2 * c + 1];
            are[tn][tm] = MF::mac(ar, lr, are[tn][tm]);
            aim[tn][tm] = MF::mac(ai, li, aim[tn][tm]);
            a3[tn][tm] = MF::mac(ar + ai, ld, a3[tn][tm]);
          }
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
#pragma unroll
        for (int tm = 0; tm < 2; ++tm) {
          const MT bre = bv[tm][j];
          const MT bim = (j & 1) ? bv[tm][j - 1] : -bv[tm][j + 1];
#pragma unroll
          for (int tn = 0; tn < 2; ++tn) {
            are[tn][tm] = MF::mac(av[tn][j], bre, are[tn][tm]);
            aim[tn][tm] = MF::mac(av[tn][j], bim, aim[tn][tm]);
          }
        }
      }
    }
    if (kPromote && (st % kFlush) == kFlush - 1) flush();
    if (st + 1 < nstage) lstore(buf ^ 1, regs);
    __syncthreads();
  }
  if (kPromote) flush();
  double2* out = part + (size_t)z * ldn * ldm;
#pragma unroll
  for (int tn = 0; tn < 2; ++tn)
#pragma unroll
    for (int tm = 0; tm < 2; ++tm)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int n = tile_n + n0 + 16 * tn + MF::row(g, v);
        const int m = tile_m + m0 + 16 * tm + i;
        if (n < nR && m < nL) {
          double2 o;
          if constexpr (kPromote) {
            o = make_double2(fre[tn][tm][v], fim[tn][tm][v]);
          } else if constexpr (k3M) {
            const double p1 = (double)are[tn][tm][v], p2 = (double)aim[tn][tm][v];
            o = make_double2(p1 + p2, ((double)a3[tn][tm][v] - p1) + p2);
          } else {
            o = make_double2((double)are[tn][tm][v], (double)aim[tn][tm][v]);
          }
          out[(size_t)n * ldm + m] = o;
        }
      }
}

// ---- 3-pass split-BF16 mode ------------------------------------------------
// Each operand real x is split as hi = bf16(x), lo = bf16(x - hi) (round to
// nearest even, from the FP32 value of x), and every product as
//   x y ~ hi_x hi_y + hi_x lo_y + lo_x hi_y
// (the lo lo term, ~2^-16 relative, is dropped) on v_mfma_f32_16x16x32_bf16:
// one instruction per 16 x 16 tile consumes a whole 32-real stage (8 reals
// per lane group), so a stage costs 6 MFMAs per tile pair (3 Re, 3 Im) where
// the FP32 mode needs 16.  FP32 accumulators are promoted into FP64 every 8
// stages, as in the FP32 mode.  The tile scheme (64 x 64 per workgroup, LDS
// double buffering, XCD-aware K-chunk dispatch) is k_contract's; the LDS holds
// the hi and lo planes as bf16 pairs.
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
// split_bf16_pair: bf16_planes.h

// L'' = (-Li, Lr) on packed (Re, Im) bf16 pairs: swap the halves, negate the new low half
__device__ __forceinline__ uint4 swap_neg(uint4 w) {
  auto f = [](uint32_t x) { return ((x >> 16) ^ 0x8000u) | (x << 16); };
  return make_uint4(f(w.x), f(w.y), f(w.z), f(w.w));
}

__device__ __forceinline__ f4 mfma_bf16(uint4 a, uint4 b, f4 c) {
  bf8 av, bv;
  __builtin_memcpy(&av, &a, 16);
  __builtin_memcpy(&bv, &b, 16);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
}

template <typename VT>
__global__ void __launch_bounds__(256)
    k_contract_bf16x3(const VT* __restrict__ L, int nL, const VT* __restrict__ R, int nR, int64_t K2,
                      int64_t P2, int64_t kchunk, int tm, int ntiles, int splits,
                      double2* __restrict__ part, int ldn, int ldm) {
  constexpr int LDB = kStep / 2 + 4;                // LDS row stride in 32-bit bf16 pairs (+16 B pad)
  constexpr int WPR = kStep * (int)sizeof(VT) / 16; // 16-byte global words per row segment
  constexpr int EPW = 16 / (int)sizeof(VT);         // VT elements per 16-byte word
  constexpr int NW = 2 * 64 * WPR / 256;            // global words per thread per stage (A+B)
  __shared__ uint32_t lhi[2][2][64 * LDB], llo[2][2][64 * LDB];

  int tile, z;
  if ((splits & 7) == 0) {
    const int b = blockIdx.x, local = b >> 3;
    tile = local % ntiles;
    z = (local / ntiles) * 8 + (b & 7);
  } else {
    tile = blockIdx.x % ntiles;
    z = blockIdx.x / ntiles;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, i = lane & 15;
  const int tile_n = (tile / tm) * 64, tile_m = (tile % tm) * 64;
  const int n0 = (wave >> 1) * 32, m0 = (wave & 1) * 32;
  const int64_t kbeg = (int64_t)z * kchunk;
  const int64_t kend = min(kbeg + kchunk, K2);
  const int nstage = kend > kbeg ? (int)((kend - kbeg) / kStep) : 0;

  const VT* src[NW];
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    const int w = tid + 256 * q;
    const int op = w / (64 * WPR), rem = w % (64 * WPR);
    const int row = rem / WPR, wi = rem % WPR;
    const int grow = op == 0 ? min(tile_n + row, nR - 1) : min(tile_m + row, nL - 1);
    src[q] = (op == 0 ? R : L) + (size_t)grow * P2 + kbeg + wi * EPW;
  }
  auto gload = [&](int stage, float4* regs) {
#pragma unroll
    for (int q = 0; q < NW; ++q) regs[q] = *reinterpret_cast<const float4*>(src[q] + (int64_t)stage * kStep);
  };
  auto lstore = [&](int buf, const float4* regs) {
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      const int w = tid + 256 * q;
      const int op = w / (64 * WPR), rem = w % (64 * WPR);
      const int row = rem / WPR, wi = rem % WPR;
      const VT* e = reinterpret_cast<const VT*>(&regs[q]);
      uint32_t* ph = &lhi[buf][op][row * LDB + wi * EPW / 2];
      uint32_t* pl = &llo[buf][op][row * LDB + wi * EPW / 2];
#pragma unroll
      for (int t = 0; t < EPW / 2; ++t) {
        uint32_t h, l;
        split_bf16_pair((float)e[2 * t], (float)e[2 * t + 1], h, l);
        ph[t] = h;
        pl[t] = l;
      }
    }
  };

  f4 are[2][2], aim[2][2];
  double fre[2][2][4], fim[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      are[a][b] = aim[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int v = 0; v < 4; ++v) fre[a][b][v] = fim[a][b][v] = 0.0;
    }
  auto flush = [&]() {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          fre[a][b][v] += (double)are[a][b][v];
          fim[a][b][v] += (double)aim[a][b][v];
        }
        are[a][b] = aim[a][b] = f4{0.f, 0.f, 0.f, 0.f};
      }
  };
  constexpr int kFlush = 8;

  float4 regs[NW];
  if (nstage > 0) {
    gload(0, regs);
    lstore(0, regs);
  }
  __syncthreads();
  for (int st = 0; st < nstage; ++st) {
    const int buf = st & 1;
    if (st + 1 < nstage) gload(st + 1, regs);
    uint4 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int ra = (n0 + 16 * t + i) * LDB + 4 * g, rb = (m0 + 16 * t + i) * LDB + 4 * g;
      ah[t] = *reinterpret_cast<const uint4*>(&lhi[buf][0][ra]);
      al[t] = *reinterpret_cast<const uint4*>(&llo[buf][0][ra]);
      bh[t] = *reinterpret_cast<const uint4*>(&lhi[buf][1][rb]);
      bl[t] = *reinterpret_cast<const uint4*>(&llo[buf][1][rb]);
    }
#pragma unroll
    for (int tmi = 0; tmi < 2; ++tmi) {
      const uint4 b2h = swap_neg(bh[tmi]), b2l = swap_neg(bl[tmi]);
#pragma unroll
      for (int tn = 0; tn < 2; ++tn) {
        are[tn][tmi] = mfma_bf16(al[tn], bh[tmi], are[tn][tmi]);
        are[tn][tmi] = mfma_bf16(ah[tn], bl[tmi], are[tn][tmi]);
        are[tn][tmi] = mfma_bf16(ah[tn], bh[tmi], are[tn][tmi]);
        aim[tn][tmi] = mfma_bf16(al[tn], b2h, aim[tn][tmi]);
        aim[tn][tmi] = mfma_bf16(ah[tn], b2l, aim[tn][tmi]);
        aim[tn][tmi] = mfma_bf16(ah[tn], b2h, aim[tn][tmi]);
      }
    }
    if ((st % kFlush) == kFlush - 1) flush();
    if (st + 1 < nstage) lstore(buf ^ 1, regs);
    __syncthreads();
  }
  flush();
  double2* out = part + (size_t)z * ldn * ldm;
#pragma unroll
  for (int tn = 0; tn < 2; ++tn)
#pragma unroll
    for (int tmi = 0; tmi < 2; ++tmi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int n = tile_n + n0 + 16 * tn + 4 * g + v;  // f32 16x16 accumulator rows
        const int m = tile_m + m0 + 16 * tmi + i;
        if (n < nR && m < nL) out[(size_t)n * ldm + m] = make_double2(fre[tn][tmi][v], fim[tn][tmi][v]);
      }
}

// Large-table form of the split-BF16 contraction: a 128 x 128 workgroup tile
// on sixteen waves of 32 x 32 (2 x 2 MFMA tiles), four waves per SIMD, so
// the VALU split of the next stage by some waves runs beside the MFMAs of
// the others (measured on the 512 x 512 table: one wave per SIMD with 64 x 64
// per wave 18.7 ms, two waves 14.5 ms, four waves 13.5 ms).  Each operand
// fragment read from LDS feeds 2 tiles.  The MFMA accumulators hold FP32 sums
// over one K-chunk of at most kChunkBF16 reals (no FP64 registers, which is
// what leaves room for four waves); the host sizes the split so that chunks
// stay within that bound, and the FP32 chunk partials are added in FP64 in a
// fixed order.
constexpr int64_t kChunkBF16 = 8192;
constexpr int kT128Threads = 1024;

template <typename VT>
__global__ void __launch_bounds__(kT128Threads)
    k_contract_bf16x3_t128(const VT* __restrict__ L, int nL, const VT* __restrict__ R, int nR,
                           int64_t K2, int64_t P2, int64_t kchunk, int tm, int ntiles, int splits,
                           float2* __restrict__ part, int ldn, int ldm) {
  constexpr int TT = 128;
  constexpr int NT = kT128Threads;
  constexpr int WN = 2, WM = 2;  // MFMA tiles per wave along n (R rows) and m (L rows)
  constexpr int LDB = kStep / 2 + 4;
  constexpr int WPR = kStep * (int)sizeof(VT) / 16;
  constexpr int EPW = 16 / (int)sizeof(VT);
  constexpr int NW = 2 * TT * WPR / NT;
  static_assert((TT / (16 * WN)) * (TT / (16 * WM)) == NT / 64, "waves cover the tile");
  __shared__ uint32_t lhi[2][2][TT * LDB], llo[2][2][TT * LDB];

  int tile, z;
  if ((splits & 7) == 0) {
    const int b = blockIdx.x, local = b >> 3;
    tile = local % ntiles;
    z = (local / ntiles) * 8 + (b & 7);
  } else {
    tile = blockIdx.x % ntiles;
    z = blockIdx.x / ntiles;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, i = lane & 15;
  const int tile_n = (tile / tm) * TT, tile_m = (tile % tm) * TT;
  constexpr int WAVES_M = TT / (16 * WM);
  const int n0 = (wave / WAVES_M) * (16 * WN), m0 = (wave % WAVES_M) * (16 * WM);
  const int64_t kbeg = (int64_t)z * kchunk;
  const int64_t kend = min(kbeg + kchunk, K2);
  const int nstage = kend > kbeg ? (int)((kend - kbeg) / kStep) : 0;

  const VT* src[NW];
#pragma unroll
  for (int q = 0; q < NW; ++q) {
    const int w = tid + NT * q;
    const int op = w / (TT * WPR), rem = w % (TT * WPR);
    const int row = rem / WPR, wi = rem % WPR;
    const int grow = op == 0 ? min(tile_n + row, nR - 1) : min(tile_m + row, nL - 1);
    src[q] = (op == 0 ? R : L) + (size_t)grow * P2 + kbeg + wi * EPW;
  }
  auto gload = [&](int stage, float4* regs) {
#pragma unroll
    for (int q = 0; q < NW; ++q) regs[q] = *reinterpret_cast<const float4*>(src[q] + (int64_t)stage * kStep);
  };
  auto lstore = [&](int buf, const float4* regs) {
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      const int w = tid + NT * q;
      const int op = w / (TT * WPR), rem = w % (TT * WPR);
      const int row = rem / WPR, wi = rem % WPR;
      const VT* e = reinterpret_cast<const VT*>(&regs[q]);
      uint32_t* ph = &lhi[buf][op][row * LDB + wi * EPW / 2];
      uint32_t* pl = &llo[buf][op][row * LDB + wi * EPW / 2];
#pragma unroll
      for (int t = 0; t < EPW / 2; ++t) {
        uint32_t h, l;
        split_bf16_pair((float)e[2 * t], (float)e[2 * t + 1], h, l);
        ph[t] = h;
        pl[t] = l;
      }
    }
  };

  f4 are[WN][WM], aim[WN][WM];
#pragma unroll
  for (int a = 0; a < WN; ++a)
#pragma unroll
    for (int b = 0; b < WM; ++b) are[a][b] = aim[a][b] = f4{0.f, 0.f, 0.f, 0.f};

  float4 regs[NW];
  if (nstage > 0) {
    gload(0, regs);
    lstore(0, regs);
  }
  __syncthreads();
  for (int st = 0; st < nstage; ++st) {
    const int buf = st & 1;
    if (st + 1 < nstage) gload(st + 1, regs);
    uint4 ah[WN], al[WN];
#pragma unroll
    for (int t = 0; t < WN; ++t) {
      const int ra = (n0 + 16 * t + i) * LDB + 4 * g;
      ah[t] = *reinterpret_cast<const uint4*>(&lhi[buf][0][ra]);
      al[t] = *reinterpret_cast<const uint4*>(&llo[buf][0][ra]);
    }
#pragma unroll
    for (int tmi = 0; tmi < WM; ++tmi) {
      const int rb = (m0 + 16 * tmi + i) * LDB + 4 * g;
      const uint4 bh = *reinterpret_cast<const uint4*>(&lhi[buf][1][rb]);
      const uint4 bl = *reinterpret_cast<const uint4*>(&llo[buf][1][rb]);
      const uint4 b2h = swap_neg(bh), b2l = swap_neg(bl);
#pragma unroll
      for (int tn = 0; tn < WN; ++tn) {
        are[tn][tmi] = mfma_bf16(al[tn], bh, are[tn][tmi]);
        are[tn][tmi] = mfma_bf16(ah[tn], bl, are[tn][tmi]);
        are[tn][tmi] = mfma_bf16(ah[tn], bh, are[tn][tmi]);
        aim[tn][tmi] = mfma_bf16(al[tn], b2h, aim[tn][tmi]);
        aim[tn][tmi] = mfma_bf16(ah[tn], b2l, aim[tn][tmi]);
        aim[tn][tmi] = mfma_bf16(ah[tn], b2h, aim[tn][tmi]);
      }
    }
    if (st + 1 < nstage) lstore(buf ^ 1, regs);
    __syncthreads();
  }
  float2* out = part + (size_t)z * ldn * ldm;
#pragma unroll
  for (int tn = 0; tn < WN; ++tn)
#pragma unroll
    for (int tmi = 0; tmi < WM; ++tmi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int n = tile_n + n0 + 16 * tn + 4 * g + v;
        const int m = tile_m + m0 + 16 * tmi + i;
        if (n < nR && m < nL) out[(size_t)n * ldm + m] = make_float2(are[tn][tmi][v], aim[tn][tmi][v]);
      }
}

// The same contraction from operands already stored as split-BF16 planes
// (bf16_planes.h), written by the recurrence passes that produce the slots.
// A stage row is then 128 contiguous bytes (the 16 hi words, then the 16 lo
// words, exactly the LDS row of the kernel above without its pad), so the
// stage is copied to LDS without a split, and the operand bytes read per
// stage are half those of FP64 slots.  The LDS layout, the fragment reads and
// the MFMA order per accumulator are those of k_contract_bf16x3_t128, so the
// partials are bitwise the same as from the FP64 (or FP32) slots the planes
// were split from.
//
// The 128 x 128 tile runs on eight waves of 32 (n) x 64 (m), so that two
// workgroups (80 KB of LDS each) share a CU: while one waits at its per-stage
// barrier for its stage loads, the other's MFMAs run.  Measured on the
// 512 x 512, K = 2.1M complex table on random planes: 10.3-10.5 ms, against
// 10.7-10.9 ms for sixteen waves of 32 x 32 in one workgroup per CU with loads
// two stages ahead, and 11.0-11.2 ms with loads one stage ahead.  (The
// register limit of four waves per SIMD leaves no room for a second set of
// stage words: two-ahead loads on this layout spilled 40 registers.)
constexpr int kPlaneThreads = 512;

__global__ void __launch_bounds__(kPlaneThreads, 4)
    k_contract_planes_t128(const uint4* __restrict__ L, int nL, const uint4* __restrict__ R, int nR,
                           int64_t K2, int64_t P2, int64_t kchunk, int tm, int ntiles, int splits,
                           float2* __restrict__ part, int ldn, int ldm) {
  constexpr int TT = 128;
  constexpr int NT = kPlaneThreads;
  constexpr int WN = 2, WM = 4;        // MFMA tiles per wave along n (R rows) and m (L rows)
  constexpr int LDB = kStep / 2 + 4;   // LDS row stride in words (16 pairs + pad)
  constexpr int PW = 8;                // 16-byte words per row and stage: 4 hi, then 4 lo
  constexpr int PNW = 2 * TT * PW / NT;
  static_assert(PNW == 4, "four plane words per thread and stage");
  static_assert((TT / (16 * WN)) * (TT / (16 * WM)) == NT / 64, "waves cover the tile");
  // [plane][buffer][operand][row * LDB]
  __shared__ uint32_t lds[2][2][2][TT * LDB];
  constexpr int PLANE = 2 * 2 * TT * LDB, BUF = 2 * TT * LDB;

  int tile, z;
  if ((splits & 7) == 0) {
    const int b = blockIdx.x, local = b >> 3;
    tile = local % ntiles;
    z = (local / ntiles) * 8 + (b & 7);
  } else {
    tile = blockIdx.x % ntiles;
    z = blockIdx.x / ntiles;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, i = lane & 15;
  const int tile_n = (tile / tm) * TT, tile_m = (tile % tm) * TT;
  constexpr int WAVES_M = TT / (16 * WM);
  const int n0 = (wave / WAVES_M) * (16 * WN), m0 = (wave % WAVES_M) * (16 * WM);
  const int64_t kbeg = (int64_t)z * kchunk;
  const int64_t kend = min(kbeg + kchunk, K2);
  const int nstage = kend > kbeg ? (int)((kend - kbeg) / kStep) : 0;

  // the four words this thread moves per stage: global source, LDS offset
  const uint4* src[PNW];
  int loff[PNW];
#pragma unroll
  for (int q = 0; q < PNW; ++q) {
    const int w = tid + NT * q;
    const int op = w / (TT * PW), rem = w % (TT * PW);
    const int row = rem / PW, wi = rem % PW;
    const int grow = op == 0 ? min(tile_n + row, nR - 1) : min(tile_m + row, nL - 1);
    // a slot of P2 / 2 complex values is P2 words = P2 / 4 16-byte words
    src[q] = (op == 0 ? R : L) + (size_t)grow * (size_t)(P2 / 4) + kbeg / 4 + wi;
    loff[q] = (wi >> 2) * PLANE + (op * TT + row) * LDB + (wi & 3) * 4;
  }
  uint32_t* const base = &lds[0][0][0][0];
  uint4 w0, w1, w2, w3;  // (named words, not an array: an array of them went to scratch)
  auto gload = [&](int stage) {
    const int64_t o = (int64_t)stage * PW;
    w0 = src[0][o];
    w1 = src[1][o];
    w2 = src[2][o];
    w3 = src[3][o];
  };
  auto lstore = [&](int buf) {
    uint32_t* p = base + buf * BUF;
    *reinterpret_cast<uint4*>(p + loff[0]) = w0;
    *reinterpret_cast<uint4*>(p + loff[1]) = w1;
    *reinterpret_cast<uint4*>(p + loff[2]) = w2;
    *reinterpret_cast<uint4*>(p + loff[3]) = w3;
  };

  f4 are[WN][WM], aim[WN][WM];
#pragma unroll
  for (int a = 0; a < WN; ++a)
#pragma unroll
    for (int b = 0; b < WM; ++b) are[a][b] = aim[a][b] = f4{0.f, 0.f, 0.f, 0.f};

  if (nstage > 0) {
    gload(0);
    lstore(0);
  }
  __syncthreads();
  for (int st = 0; st < nstage; ++st) {
    const int buf = st & 1;
    if (st + 1 < nstage) gload(st + 1);
    const uint32_t* p = base + buf * BUF;
    uint4 ah[WN], al[WN];
#pragma unroll
    for (int t = 0; t < WN; ++t) {
      const int ra = (n0 + 16 * t + i) * LDB + 4 * g;
      ah[t] = *reinterpret_cast<const uint4*>(p + ra);
      al[t] = *reinterpret_cast<const uint4*>(p + PLANE + ra);
    }
#pragma unroll
    for (int tmi = 0; tmi < WM; ++tmi) {
      const int rb = (TT + m0 + 16 * tmi + i) * LDB + 4 * g;
      const uint4 bh = *reinterpret_cast<const uint4*>(p + rb);
      const uint4 bl = *reinterpret_cast<const uint4*>(p + PLANE + rb);
      const uint4 b2h = swap_neg(bh), b2l = swap_neg(bl);
#pragma unroll
      for (int tn = 0; tn < WN; ++tn) {
        are[tn][tmi] = mfma_bf16(al[tn], bh, are[tn][tmi]);
        are[tn][tmi] = mfma_bf16(ah[tn], bl, are[tn][tmi]);
        are[tn][tmi] = mfma_bf16(ah[tn], bh, are[tn][tmi]);
        aim[tn][tmi] = mfma_bf16(al[tn], b2h, aim[tn][tmi]);
        aim[tn][tmi] = mfma_bf16(ah[tn], b2l, aim[tn][tmi]);
        aim[tn][tmi] = mfma_bf16(ah[tn], b2h, aim[tn][tmi]);
      }
    }
    if (st + 1 < nstage) lstore(buf ^ 1);
    __syncthreads();
  }
  float2* out = part + (size_t)z * ldn * ldm;
#pragma unroll
  for (int tn = 0; tn < WN; ++tn)
#pragma unroll
    for (int tmi = 0; tmi < WM; ++tmi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int n = tile_n + n0 + 16 * tn + 4 * g + v;
        const int m = tile_m + m0 + 16 * tmi + i;
        if (n < nR && m < nL) out[(size_t)n * ldm + m] = make_float2(are[tn][tmi][v], aim[tn][tmi][v]);
      }
}

// FP32 chunk partials -> FP64 sums in chunk order, then scale / mask / accumulate
__global__ void k_contract_reduce_f32(const float2* __restrict__ part, int splits, int nR, int nL,
                                      int ldn, int ldm, double scale, double2* __restrict__ mu,
                                      int64_t NC, int64_t n0, int64_t m0, int parity) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nR * nL) return;
  const int n = (int)(idx / nL);
  const int m = (int)(idx % nL);
  const int64_t s = n0 + n + m0 + m + 2;
  const bool keep = parity == KPM_PARITY_NONE || (parity == KPM_PARITY_ODD && (s & 1)) ||
                    (parity == KPM_PARITY_EVEN && !(s & 1));
  if (!keep) return;
  double sr = 0.0, si = 0.0;
  for (int z = 0; z < splits; ++z) {
    const float2 v = part[((size_t)z * ldn + n) * ldm + m];
    sr += (double)v.x;
    si += (double)v.y;
  }
  double2* dst = mu + (n0 + n) + NC * (m0 + m);
  double2 cur = *dst;
  cur.x += scale * sr;
  cur.y += scale * si;
  *dst = cur;
}

__global__ void k_contract_reduce(const double2* __restrict__ part, int splits, int nR, int nL,
                                  int ldn, int ldm, double scale, double2* __restrict__ mu,
                                  int64_t NC, int64_t n0, int64_t m0, int parity) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nR * nL) return;
  const int n = (int)(idx / nL);
  const int m = (int)(idx % nL);
  // reference indices are 1-based: n_ref + m_ref = (n0+n+1) + (m0+m+1)
  const int64_t s = n0 + n + m0 + m + 2;
  const bool keep = parity == KPM_PARITY_NONE || (parity == KPM_PARITY_ODD && (s & 1)) ||
                    (parity == KPM_PARITY_EVEN && !(s & 1));
  if (!keep) return;
  double sr = 0.0, si = 0.0;
  for (int z = 0; z < splits; ++z) {
    const double2 v = part[((size_t)z * ldn + n) * ldm + m];
    sr += v.x;
    si += v.y;
  }
  double2* dst = mu + (n0 + n) + NC * (m0 + m);
  double2 cur = *dst;
  cur.x += scale * sr;
  cur.y += scale * si;
  *dst = cur;
}

// prec < 0: the slots are split-BF16 planes (k_contract_planes_t128)
static int launch_contract_bf16_t128(kpm_ctx* ctx, const void* Lslots, int nL, const void* Rslots,
                                     int nR, int64_t K2, int64_t P2, int prec, double scale,
                                     void* mu_dev, int64_t NC, int64_t n0, int64_t m0, int parity) {
  const int tn = (nR + 127) / 128, tm = (nL + 127) / 128;
  const int ntiles = tn * tm;
  const int64_t steps = K2 / kStep;
  // chunks of at most kChunkBF16 reals (FP32 accumulation length), at least
  // 4 workgroups per CU, and partials within 2 GiB
  int64_t splits = std::max<int64_t>((K2 + kChunkBF16 - 1) / kChunkBF16,
                                     (int64_t)(4 * std::max(ctx->num_cus, 1)) / ntiles);
  const int64_t cap = std::max<int64_t>(1, ((int64_t)2 << 30) / ((int64_t)nR * nL * (int64_t)sizeof(float2)));
  splits = std::min<int64_t>(std::min(splits, cap), std::max<int64_t>(1, steps));
  if (splits >= 8) splits = std::min<int64_t>((splits + 7) / 8 * 8, std::max<int64_t>(8, cap / 8 * 8));
  const int64_t steps_per = (steps + splits - 1) / splits;
  const int64_t kchunk = steps_per * kStep;
  if ((int64_t)ntiles * splits > INT32_MAX) return fail(KPM_EINVAL, "contraction grid too large");
  const int ldn = nR, ldm = nL;
  KPM_TRY(ensure(ctx, ctx->gemm_part, (size_t)splits * ldn * ldm * sizeof(float2)));
  const dim3 grid((unsigned)(ntiles * splits));
  const int sp = (int)splits;
  if (prec < 0)
    hipLaunchKernelGGL(k_contract_planes_t128, grid, dim3(kPlaneThreads), 0, ctx->stream,
                       (const uint4*)Lslots, nL, (const uint4*)Rslots, nR, K2, P2, kchunk, tm, ntiles, sp,
                       (float2*)ctx->gemm_part.ptr, ldn, ldm);
  else if (prec == KPM_VEC_F64)
    hipLaunchKernelGGL((k_contract_bf16x3_t128<double>), grid, dim3(kT128Threads), 0, ctx->stream,
                       (const double*)Lslots, nL, (const double*)Rslots, nR, K2, P2, kchunk, tm, ntiles, sp,
                       (float2*)ctx->gemm_part.ptr, ldn, ldm);
  else
    hipLaunchKernelGGL((k_contract_bf16x3_t128<float>), grid, dim3(kT128Threads), 0, ctx->stream,
                       (const float*)Lslots, nL, (const float*)Rslots, nR, K2, P2, kchunk, tm, ntiles, sp,
                       (float2*)ctx->gemm_part.ptr, ldn, ldm);
  KPM_HIP_TRY(hipGetLastError());
  const int64_t total = (int64_t)nR * nL;
  hipLaunchKernelGGL(k_contract_reduce_f32, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     ctx->stream, (const float2*)ctx->gemm_part.ptr, sp, nR, nL, ldn, ldm, scale,
                     (double2*)mu_dev, NC, n0, m0, parity);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

int launch_contract_planes(kpm_ctx* ctx, const void* Lplanes, int nL, const void* Rplanes, int nR,
                           int64_t K, int64_t pitch, double scale, void* mu_dev, int64_t NC, int64_t n0,
                           int64_t m0, int parity) {
  if ((2 * K) % kStep != 0) return fail(KPM_EINVAL, "contraction slot length must be a multiple of 16");
  if (pitch < K || pitch % 8 != 0) return fail(KPM_EINVAL, "slot pitch must be >= K and a multiple of 8");
  if (nL < 128 || nR < 128) return fail(KPM_EINVAL, "plane contraction needs 128 or more slots a side");
  return launch_contract_bf16_t128(ctx, Lplanes, nL, Rplanes, nR, 2 * K, 2 * pitch, -1, scale, mu_dev, NC,
                                   n0, m0, parity);
}

int launch_contract(kpm_ctx* ctx, const void* Lslots, int nL, const void* Rslots, int nR,
                    int64_t K, int64_t pitch, int prec, int gemm_mode, double scale, void* mu_dev,
                    int64_t NC, int64_t n0, int64_t m0, int parity) {
  const int64_t K2 = 2 * K;      // reals contracted per slot (a multiple of kStep)
  const int64_t P2 = 2 * pitch;  // reals between consecutive slots
  if (K2 % kStep != 0) return fail(KPM_EINVAL, "contraction slot length must be a multiple of 16");
  if (pitch < K || pitch % 8 != 0) return fail(KPM_EINVAL, "slot pitch must be >= K and a multiple of 8");
  const int tn = (nR + 63) / 64;
  const int tm = (nL + 63) / 64;
  const int ntiles = tn * tm;
  const int64_t steps = K2 / kStep;
  int64_t splits = std::max<int64_t>(1, (int64_t)(4 * std::max(ctx->num_cus, 1)) / ntiles);
  splits = std::min<int64_t>(splits, std::max<int64_t>(1, steps / 16));
  splits = std::min<int64_t>(splits, 1024);
  if (splits >= 8) splits = (splits + 7) / 8 * 8;  // whole K-chunk groups per XCD
  const int64_t steps_per = (steps + splits - 1) / splits;
  const int64_t kchunk = steps_per * kStep;  // trailing chunks may be empty (zero partials)
  if ((int64_t)ntiles * splits > INT32_MAX) return fail(KPM_EINVAL, "contraction grid too large");
  const int ldn = nR, ldm = nL;
  KPM_TRY(ensure(ctx, ctx->gemm_part, (size_t)splits * ldn * ldm * sizeof(double2)));
  const dim3 grid((unsigned)(ntiles * splits));
  const int sp = (int)splits;
  if (gemm_mode == KPM_GEMM_BF16X3 && nL >= 128 && nR >= 128)
    return launch_contract_bf16_t128(ctx, Lslots, nL, Rslots, nR, K2, P2, prec, scale, mu_dev, NC, n0,
                                     m0, parity);
  if (gemm_mode == KPM_GEMM_BF16X3) {
    if (prec == KPM_VEC_F64)
      hipLaunchKernelGGL((k_contract_bf16x3<double>), grid, dim3(256), 0, ctx->stream,
                         (const double*)Lslots, nL, (const double*)Rslots, nR, K2, P2, kchunk, tm,
                         ntiles, sp, (double2*)ctx->gemm_part.ptr, ldn, ldm);
    else
      hipLaunchKernelGGL((k_contract_bf16x3<float>), grid, dim3(256), 0, ctx->stream,
                         (const float*)Lslots, nL, (const float*)Rslots, nR, K2, P2, kchunk, tm,
                         ntiles, sp, (double2*)ctx->gemm_part.ptr, ldn, ldm);
  } else if (prec == KPM_VEC_F64) {
    if (gemm_mode == KPM_GEMM_F32)
      hipLaunchKernelGGL((k_contract<double, float>), grid, dim3(256), 0, ctx->stream,
                         (const double*)Lslots, nL, (const double*)Rslots, nR, K2, P2, kchunk, tm,
                         ntiles, sp, (double2*)ctx->gemm_part.ptr, ldn, ldm);
    else
      hipLaunchKernelGGL((k_contract<double, double>), grid, dim3(256), 0, ctx->stream,
                         (const double*)Lslots, nL, (const double*)Rslots, nR, K2, P2, kchunk, tm,
                         ntiles, sp, (double2*)ctx->gemm_part.ptr, ldn, ldm);
  } else {
    if (gemm_mode == KPM_GEMM_F32)
      hipLaunchKernelGGL((k_contract<float, float>), grid, dim3(256), 0, ctx->stream,
                         (const float*)Lslots, nL, (const float*)Rslots, nR, K2, P2, kchunk, tm,
                         ntiles, sp, (double2*)ctx->gemm_part.ptr, ldn, ldm);
    else
      hipLaunchKernelGGL((k_contract<float, double>), grid, dim3(256), 0, ctx->stream,
                         (const float*)Lslots, nL, (const float*)Rslots, nR, K2, P2, kchunk, tm,
                         ntiles, sp, (double2*)ctx->gemm_part.ptr, ldn, ldm);
  }
  KPM_HIP_TRY(hipGetLastError());
  const int64_t total = (int64_t)nR * nL;
  hipLaunchKernelGGL(k_contract_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     ctx->stream, (const double2*)ctx->gemm_part.ptr, (int)splits, nR, nL, ldn,
                     ldm, scale, (double2*)mu_dev, NC, n0, m0, parity);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

// pseudo-random values in [-0.5, 0.5) from a 64-bit mix of the index (the
// splitmix64 finaliser): synthetic contraction operands
__device__ inline double hash_unit(uint64_t i, uint64_t seed) {
  uint64_t x = i * 0x9E3779B97F4A7C15ull + seed;
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (double)(x >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

template <typename E>
__global__ void k_fill_unit(E* __restrict__ p, int64_t n, uint64_t seed) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (E)hash_unit((uint64_t)i, seed);
}

// planes of nslots slots of K complex values, pitch complex values apart
__global__ void k_fill_unit_planes(uint32_t* __restrict__ p, int64_t nslots, int64_t K, int64_t pitch,
                                   uint64_t seed) {
  const int64_t n = nslots * K;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t slot = i / K, k = i - slot * K;
    store_plane_value(p + slot * 2 * pitch, k, hash_unit(2 * (uint64_t)i, seed),
                      hash_unit(2 * (uint64_t)i + 1, seed));
  }
}

static int fill_operand(kpm_ctx* ctx, void* p, int64_t nslots, int64_t K, int64_t pitch, int vec_prec,
                        uint64_t seed) {
  const dim3 grid((unsigned)std::max(ctx->num_cus, 8) * 16), block(256);
  if (vec_prec == KPM_VEC_PLANES)
    hipLaunchKernelGGL(k_fill_unit_planes, grid, block, 0, ctx->stream, (uint32_t*)p, nslots, K, pitch, seed);
  else if (vec_prec == KPM_VEC_F32)
    hipLaunchKernelGGL(k_fill_unit<float>, grid, block, 0, ctx->stream, (float*)p, nslots * pitch * 2, seed);
  else
    hipLaunchKernelGGL(k_fill_unit<double>, grid, block, 0, ctx->stream, (double*)p, nslots * pitch * 2, seed);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_bench_contract(kpm_ctx* ctx, int64_t nL, int64_t nR, int64_t K, int64_t pitch, int vec_prec,
                       int gemm_mode, int64_t reps, float* ms) {
  if (!ctx || !ms || nL < 1 || nR < 1 || reps < 1 || K < 16 || K % 16 != 0)
    return fail(KPM_EINVAL, "bad arguments (K must be a positive multiple of 16)");
  const bool planes = vec_prec == KPM_VEC_PLANES;
  if (pitch == 0) pitch = K + 256 / elem_bytes(vec_prec == KPM_VEC_F32 ? KPM_VEC_F32 : KPM_VEC_F64);
  if (pitch < K || pitch % 8 != 0) return fail(KPM_EINVAL, "pitch must be >= K and a multiple of 8");
  if (vec_prec != KPM_VEC_F64 && vec_prec != KPM_VEC_F32 && !planes) return fail(KPM_EINVAL, "bad vec_prec");
  if (gemm_mode != KPM_GEMM_F64 && gemm_mode != KPM_GEMM_F32 && gemm_mode != KPM_GEMM_BF16X3)
    return fail(KPM_EINVAL, "bad gemm_mode");
  if (planes && (gemm_mode != KPM_GEMM_BF16X3 || nL < 128 || nR < 128))
    return fail(KPM_EINVAL, "plane slots: split-BF16 mode, 128 or more slots a side");
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  const size_t eb = planes ? 8 : (size_t)elem_bytes(vec_prec);  // bytes per complex value
  const int64_t NC = std::max(nL, nR);
  Buffer lb, rb, mb;
  int rc = ensure(ctx, lb, (size_t)(nL * pitch) * eb);
  if (rc == KPM_OK) rc = ensure(ctx, rb, (size_t)(nR * pitch) * eb);
  if (rc == KPM_OK) rc = ensure(ctx, mb, (size_t)(NC * NC) * sizeof(double2));
  if (rc == KPM_OK) rc = fill_operand(ctx, lb.ptr, nL, K, pitch, vec_prec, 1);
  if (rc == KPM_OK) rc = fill_operand(ctx, rb.ptr, nR, K, pitch, vec_prec, 2);
  if (rc == KPM_OK && hipMemsetAsync(mb.ptr, 0, mb.bytes, ctx->stream) != hipSuccess)
    rc = fail(KPM_EHIP, "workspace clear failed");
  auto one = [&]() {
    return planes ? launch_contract_planes(ctx, lb.ptr, (int)nL, rb.ptr, (int)nR, K, pitch, 1.0, mb.ptr, NC,
                                           0, 0, KPM_PARITY_NONE)
                  : launch_contract(ctx, lb.ptr, (int)nL, rb.ptr, (int)nR, K, pitch, vec_prec, gemm_mode,
                                    1.0, mb.ptr, NC, 0, 0, KPM_PARITY_NONE);
  };
  if (rc == KPM_OK) rc = one();  // warm-up
  if (rc == KPM_OK && hipEventRecord(ctx->events[14], ctx->stream) != hipSuccess)
    rc = fail(KPM_EHIP, "event record failed");
  for (int64_t r = 0; r < reps && rc == KPM_OK; ++r) rc = one();
  if (rc == KPM_OK && (hipEventRecord(ctx->events[15], ctx->stream) != hipSuccess ||
                       hipEventSynchronize(ctx->events[15]) != hipSuccess ||
                       hipEventElapsedTime(ms, ctx->events[14], ctx->events[15]) != hipSuccess))
    rc = fail(KPM_EHIP, "event timing failed");
  (void)hipStreamSynchronize(ctx->stream);
  for (Buffer* b : {&lb, &rb, &mb})
    if (b->ptr) (void)hipFree(b->ptr);
  return rc;
}

}  // extern "C"
