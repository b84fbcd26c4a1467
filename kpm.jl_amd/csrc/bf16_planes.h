// bf16_planes.h -- the hi/lo BF16 split of the 3-pass split-BF16 contraction
// and the plane layout its operands can be stored in.
//
// Each real x is split as hi = bf16(x), lo = bf16(x - hi), both rounded to
// nearest even from the FP32 value of x (the residual x - hi is exact in
// FP32).  A complex value (re, im) gives a hi word (bf16 re | bf16 im << 16)
// and a lo word of the same form.
//
// Plane layout of one slot (a flat vector of K complex values, K a multiple
// of 16): block b of 16 consecutive complex values takes 32 words, the 16 hi
// words of its values in order, then their 16 lo words.  So one 32-real stage
// of the contraction is 128 contiguous bytes per slot, in the order its LDS
// tile holds them, and the slot is 8 bytes per complex value instead of 16
// (FP64).  A slot of pitch P complex values takes 2 P words.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace kpm {

typedef float bfp_f2v __attribute__((ext_vector_type(2)));
typedef __bf16 bfp_b2v __attribute__((ext_vector_type(2)));

// (hi, lo) bf16 splits of the pair (x0, x1), packed x0 in the low half: the
// FP32 -> BF16 conversions are gfx950's packed v_cvt_pk_bf16_f32 (round to
// nearest even)
__device__ __forceinline__ void split_bf16_pair(float x0, float x1, uint32_t& hi, uint32_t& lo) {
  hi = __builtin_bit_cast(uint32_t, __builtin_convertvector((bfp_f2v){x0, x1}, bfp_b2v));
  const float r0 = x0 - __uint_as_float(hi << 16);
  const float r1 = x1 - __uint_as_float(hi & 0xFFFF0000u);
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector((bfp_f2v){r0, r1}, bfp_b2v));
}

// the planes of complex value k = (re, im) of a slot (streaming stores: a
// slot is read back only by the contraction, after many other slots)
__device__ __forceinline__ void store_plane_value(uint32_t* __restrict__ planes, int64_t k, double re,
                                                  double im) {
  uint32_t hi, lo;
  split_bf16_pair((float)re, (float)im, hi, lo);
  uint32_t* p = planes + (k >> 4) * 32 + (k & 15);
  __builtin_nontemporal_store(hi, p);
  __builtin_nontemporal_store(lo, p + 16);
}

}  // namespace kpm
