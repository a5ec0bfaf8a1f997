// Hash-grid lookups shared by the encode kernels and the fused relight
// kernel (one definition, so both produce the same FP32 features).
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "kernels.hpp"

namespace nwprt {
namespace dev {

// Spatial hash of Instant-NGP (Mueller et al. 2022): x*1 ^ y*2654435761 ^
// z*805459861 in uint32 arithmetic.
__device__ __forceinline__ uint32_t grid_hash(uint32_t x, uint32_t y, uint32_t z) {
  return x ^ (y * 2654435761u) ^ (z * 805459861u);
}

// Trilinear hash-grid features of level l at position x, added to acc[0..F).
// Position pos = fmaf(x, scale_l, 0.5f) in FP32; corner c takes bit 0 / 1 / 2
// for +x / +y / +z; dense levels index x + y res + z res^2 (mod size), hashed
// ones grid_hash & (size - 1).  idx (may be null) receives the 8 indices.
template <int F, bool HF16>
__device__ __forceinline__ void accumulate_level(const nwprt_hashgrid_levels &lv, const void *hash,
                                                 int l, const float (&x)[3], float *acc,
                                                 uint32_t *idx_out) {
  const float scale = lv.scale[l];
  const uint32_t res = lv.resolution[l], size = lv.size[l];
  const bool dense = lv.dense[l] != 0;
  const uint32_t off = lv.offset[l];
  uint32_t gi[3];
  float fr[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float pos = fmaf(x[d], scale, 0.5f);
    const float fl = floorf(pos);
    gi[d] = (uint32_t)(int)fl;
    fr[d] = pos - fl;
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t cx = gi[0] + (c & 1), cy = gi[1] + ((c >> 1) & 1), cz = gi[2] + ((c >> 2) & 1);
    uint32_t idx;
    if (dense) {
      idx = cx + cy * res + cz * res * res;
      if (idx >= size) idx %= size;
    } else {
      idx = grid_hash(cx, cy, cz) & (size - 1u);
    }
    const float w = ((c & 1) ? fr[0] : 1.0f - fr[0]) * ((c & 2) ? fr[1] : 1.0f - fr[1]) *
                    ((c & 4) ? fr[2] : 1.0f - fr[2]);
    const size_t e = ((size_t)off + idx) * F;
    if (HF16) {
      const uint16_t *h = (const uint16_t *)hash + e;
      if (F == 2) {
        const uint32_t two = *(const uint32_t *)h;
        acc[0] = fmaf(w, dev::f16_to_f32((uint16_t)(two & 0xffffu)), acc[0]);
        acc[1] = fmaf(w, dev::f16_to_f32((uint16_t)(two >> 16)), acc[1]);
      } else {
#pragma unroll
        for (int j = 0; j < F; ++j) acc[j] = fmaf(w, dev::f16_to_f32(h[j]), acc[j]);
      }
    } else {
      const float *h = (const float *)hash + e;
      if (F == 2) {
        const float2 two = *(const float2 *)h;
        acc[0] = fmaf(w, two.x, acc[0]);
        acc[1] = fmaf(w, two.y, acc[1]);
      } else {
#pragma unroll
        for (int j = 0; j < F; ++j) acc[j] = fmaf(w, h[j], acc[j]);
      }
    }
    if (idx_out) idx_out[c] = idx;
  }
}

}  // namespace dev
}  // namespace nwprt
