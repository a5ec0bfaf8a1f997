// Device helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tonemap_table.hpp"

namespace nwprt {
namespace dev {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Round-to-nearest-even packing of two floats into one dword.
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
  const __bf16 a = (__bf16)lo, b = (__bf16)hi;
  return (uint32_t)__builtin_bit_cast(uint16_t, a) |
         ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}
__device__ __forceinline__ uint32_t pack_f16(float lo, float hi) {
  const _Float16 a = (_Float16)lo, b = (_Float16)hi;
  return (uint32_t)__builtin_bit_cast(uint16_t, a) |
         ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}
// pack_f16 of values that come out of an FMA chain, rounded the way the
// code reads: first to FP32, then to FP16.  The empty asm keeps the
// compiler from fusing the last FMA and the conversion into one
// v_fma_mixlo/hi_f16, which rounds the exact result once and so differs
// in rare cases -- and would make the same features depend on which kernel
// computed them.  (Not for MFMA results: their reads need the compiler's
// hazard tracking, which sees only real instructions.)
__device__ __forceinline__ uint32_t pack_f16_rounded(float lo, float hi) {
  asm volatile("" : "+v"(lo), "+v"(hi));
  return pack_f16(lo, hi);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}
__device__ __forceinline__ float f16_to_f32(uint16_t h) {
  return (float)__builtin_bit_cast(_Float16, h);
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));

// ReLU of two floats rounded (RNE) to a packed bf16 / f16 pair: one
// v_cvt_pk_*_f32 and one v_pk_max_i16.  Rounding commutes with ReLU, and a
// negative bf16 / f16 value (sign bit set) is a negative int16, so an int16
// max with 0 clears it to +0 without the NaN canonicalisation an FP max
// needs.
template <typename E>
__device__ __forceinline__ uint32_t cvt_relu2(float a, float b) {
  const f32x2 v = {a, b};
  uint32_t r;
  if constexpr (sizeof(E) == 2 && __is_same(E, __bf16))
    r = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
  else
    r = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, f16x2));
  const i16x2 m = __builtin_elementwise_max(__builtin_bit_cast(i16x2, r), (i16x2){0, 0});
  return __builtin_bit_cast(uint32_t, m);
}

// log(1 + e^x) on the hardware exp2 / log2: ln 2 * log2(1 + 2^(x log2 e)),
// and x itself above 15 (where the two agree to FP32 precision).  Below about
// -17, 1 + e^x rounds to 1 and the result is 0 instead of e^x < 2^-24.
__device__ __forceinline__ float softplus(float x) {
  const float e = __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
  const float sp = __builtin_amdgcn_logf(1.0f + e) * 0.693147180559945309f;
  return x > 15.0f ? x : sp;
}

// max(x, 0) as an integer max on the bit pattern: negative floats are
// negative integers, so one v_max_i32 and no NaN canonicalisation.
__device__ __forceinline__ float relu_f32(float x) {
  return __int_as_float(max(__float_as_int(x), 0));
}

// Neumaier-compensated accumulation: (s + c) carries the running sum.
__device__ __forceinline__ void neumaier_add(float &s, float &c, float x) {
  const float t = s + x;
  c += (fabsf(s) >= fabsf(x)) ? ((s - t) + x) : ((x - t) + s);
  s = t;
}

// Tone-map byte of one channel (Renderer/toneMap.cu:58-70): the number of
// thresholds T_b <= x, where T_b is the smallest FP32 x whose curve value
// 1.055 (x + 1e-6)^(1/2.4) - 0.055, taken in FP64, reaches b / 255.9 --
// i.e. (uint)(255.9 * clamp01(curve)) without the rounding of a float pow.
// Eight steps of binary search over the table; NaN and negative inputs give
// 0 (every comparison is false), +inf gives 255.
__device__ __forceinline__ uint32_t tonemap_byte(float x) {
  uint32_t b = 0;
#pragma unroll
  for (uint32_t step = 128; step >= 1; step >>= 1)
    if (x >= kTonemapThresholds[b + step]) b += step;
  return b;
}

// sRGB-style tone map of one pixel: R | G << 8 | B << 16 | 255 << 24.
__device__ __forceinline__ uint32_t tonemap_rgba8(float r, float g, float b) {
  return tonemap_byte(r) | (tonemap_byte(g) << 8) | (tonemap_byte(b) << 16) | (255u << 24);
}

// Pointer typed as global memory.  Device buffers that reach a kernel
// through memory (a batch's job table) cannot be told from LDS by the
// compiler, which then emits flat loads; a flat load
// counts against both the vector-memory and the LDS / scalar counters, so
// each wait for an LDS read or a scalar load then also waits for the tile
// prefetches in flight.  Loads through a gptr are global loads.
template <typename T>
using gptr = const __attribute__((address_space(1))) T *;

// Load of memory no thread writes while the kernel runs (the selected
// wavelets' light coefficients), through the constant address space, so a
// wave-uniform address is served by a scalar load.  Without it a kernel
// that loops over tiles and stores outputs in between cannot prove the
// memory unchanged, issues vector loads, and their vmcnt waits also wait for
// the tile prefetches in flight.
template <typename T>
__device__ __forceinline__ T load_invariant(const T *p) {
  return *(const __attribute__((address_space(4))) T *)p;
}

// Workgroups are dealt to the 8 XCDs round-robin; this remaps workgroup
// `bid` of `nb` so that each XCD receives a contiguous run of tiles (a
// bijection of [0, nb)).  Neighbouring pixels then share their coarse
// hash-grid entries in one XCD's L2 instead of fetching them into eight.
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int x = bid & 7, local = bid >> 3;
  return x * q + (x < r ? x : r) + local;
}

}  // namespace dev
}  // namespace nwprt
