// Device helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}
__device__ __forceinline__ float f16_to_f32(uint16_t h) {
  return (float)__builtin_bit_cast(_Float16, h);
}

// log(1 + e^x), written as max(x, 0) + log(1 + e^-|x|) so it neither
// overflows nor loses the small tail.
__device__ __forceinline__ float softplus(float x) {
  return fmaxf(x, 0.0f) + __logf(1.0f + __expf(-fabsf(x)));
}

// Neumaier-compensated accumulation: (s + c) carries the running sum.
__device__ __forceinline__ void neumaier_add(float &s, float &c, float x) {
  const float t = s + x;
  c += (fabsf(s) >= fabsf(x)) ? ((s - t) + x) : ((x - t) + s);
  s = t;
}

// sRGB-style tone map of Renderer/toneMap.cu:58-70 for one pixel:
// y = clamp01(1.055 * (x + 1e-6)^(1/2.4) - 0.055), byte = (uint)(y * 255.9),
// packed R | G << 8 | B << 16 | 255 << 24.  The multiply-subtract is one
// fused multiply-add, as a CUDA build with default contraction emits.
__device__ __forceinline__ uint32_t tonemap_rgba8(float r, float g, float b) {
  const float coeff = 1.055f, inv_exp = 1.0f / 2.4f, off = coeff - 1.0f;
  const float in[3] = {r, g, b};
  uint32_t out = 255u << 24;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float y = fmaf(coeff, powf(in[c] + 1e-6f, inv_exp), -off);
    y = fminf(1.0f, fmaxf(0.0f, y));  // fmaxf maps NaN to 0
    out |= ((uint32_t)(y * 255.9f)) << (8 * c);
  }
  return out;
}

}  // namespace dev
}  // namespace nwprt
