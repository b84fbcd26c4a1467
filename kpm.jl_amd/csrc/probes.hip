// probes.hip -- random-phase probe blocks generated on the device.
//
// psi = exp(2 pi i U) with every column scaled to unit norm and no
// centering: random_phase_vectors / normalize_by_col (src/frontend.jl:212-216,
// src/utils/vectors.jl:14-20), which kpm_1d! draws on the device for its
// random-probe form (src/moment.jl:493-497).
//
// U is the build's lane-parallel xoshiro256++ stream (not Julia's Xoshiro):
// lane j is seeded with splitmix64 outputs 4j..4j+3 of the stream started at
// `seed`, and value i of the column-major NH x NR fill comes from lane
// i % lanes at step i / lanes, U = (x >> 11) * 2^-53.  The integer stream is
// the host generator's (kpmhip.rng.Xoshiro256pp) bit for bit; the phases
// and the normalisation agree with the host to rounding.
#include <algorithm>
#include <cmath>

#include "step_kernel.h"

namespace kpm {

__device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// One thread per generator lane; lane j writes values j, j + lanes, ...
template <int PREC>
__global__ void __launch_bounds__(256)
    k_phase(typename Word<PREC>::T* __restrict__ blk, int64_t NH, int64_t NR, int64_t NRp,
            uint64_t seed, int64_t lanes) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= lanes) return;
  uint64_t s0 = splitmix_at(seed, 4 * j), s1 = splitmix_at(seed, 4 * j + 1),
           s2 = splitmix_at(seed, 4 * j + 2), s3 = splitmix_at(seed, 4 * j + 3);
  const int64_t total = NH * NR;
  for (int64_t i = j; i < total; i += lanes) {
    const uint64_t r = rotl64(s0 + s3, 23) + s0;
    const uint64_t t = s1 << 17;
    s2 ^= s0;
    s3 ^= s1;
    s1 ^= s2;
    s0 ^= s3;
    s2 ^= t;
    s3 = rotl64(s3, 45);
    const double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
    double sn, cs;
    sincos(u * 6.283185307179586, &sn, &cs);
    const int64_t col = i / NH, row = i - col * NH;
    if constexpr (PREC == KPM_VEC_F64) {
      reinterpret_cast<double2*>(blk)[row * NRp + col] = make_double2(cs, sn);
    } else {
      reinterpret_cast<float2*>(blk)[row * NRp + col] = make_float2((float)cs, (float)sn);
    }
  }
}

// column c (< NR) of every row scaled by 1/sqrt(nrm2[c]) (nrm2: NR complex, real part used)
template <int PREC>
__global__ void k_colscale(typename Word<PREC>::T* __restrict__ blk, const double2* __restrict__ nrm2,
                           int64_t NH, int64_t NR, int64_t NRp) {
  const int64_t total = NH * NR;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = idx / NR, col = idx - row * NR;
    const double s = 1.0 / sqrt(nrm2[col].x);
    if constexpr (PREC == KPM_VEC_F64) {
      double2* p = reinterpret_cast<double2*>(blk) + row * NRp + col;
      const double2 v = *p;
      *p = make_double2(v.x * s, v.y * s);
    } else {
      float2* p = reinterpret_cast<float2*>(blk) + row * NRp + col;
      const float2 v = *p;
      *p = make_float2((float)((double)v.x * s), (float)((double)v.y * s));
    }
  }
}

}  // namespace kpm

using namespace kpm;

extern "C" {

int kpm_block_random_phase(kpm_block* b, uint64_t seed, int64_t lanes) {
  if (!b) return fail(KPM_EINVAL, "null block");
  if (lanes < 1 || lanes > (int64_t)1 << 24) return fail(KPM_EINVAL, "lanes must be in [1, 2^24]");
  kpm_ctx* ctx = b->ctx;
  KPM_HIP_TRY(hipSetDevice(ctx->device));
  ++ctx->epoch;
  // padded columns stay zero
  KPM_HIP_TRY(hipMemsetAsync(b->d, 0, (size_t)(b->NH * b->NRp) * elem_bytes(b->prec), ctx->stream));
  const unsigned g = (unsigned)((lanes + 255) / 256);
  if (b->prec == KPM_VEC_F32)
    hipLaunchKernelGGL(k_phase<KPM_VEC_F32>, dim3(g), dim3(256), 0, ctx->stream,
                       (float4*)b->d, b->NH, b->NR, b->NRp, seed, lanes);
  else
    hipLaunchKernelGGL(k_phase<KPM_VEC_F64>, dim3(g), dim3(256), 0, ctx->stream,
                       (double2*)b->d, b->NH, b->NR, b->NRp, seed, lanes);
  KPM_HIP_TRY(hipGetLastError());
  int G = 0;
  KPM_TRY(launch_dot(ctx, b, b, &G));
  KPM_TRY(ensure(ctx, ctx->mu, (size_t)b->NR * sizeof(double2)));
  Combine c{1, {0, 0}, {1.0, 0.0}, {-1, -1}, {0.0, 0.0}};
  KPM_TRY(launch_combine(ctx, G, b->NR, b->NRp, c, ctx->mu.ptr, 1));
  const int64_t total = b->NH * b->NR;
  const unsigned gs = (unsigned)std::min<int64_t>(std::max<int64_t>((total + 255) / 256, 1), 65536);
  if (b->prec == KPM_VEC_F32)
    hipLaunchKernelGGL(k_colscale<KPM_VEC_F32>, dim3(gs), dim3(256), 0, ctx->stream, (float4*)b->d,
                       (const double2*)ctx->mu.ptr, b->NH, b->NR, b->NRp);
  else
    hipLaunchKernelGGL(k_colscale<KPM_VEC_F64>, dim3(gs), dim3(256), 0, ctx->stream, (double2*)b->d,
                       (const double2*)ctx->mu.ptr, b->NH, b->NR, b->NRp);
  KPM_HIP_TRY(hipGetLastError());
  return KPM_OK;
}

}  // extern "C"
