// Hash-grid lookups shared by the encode kernels and the fused relight
// kernel (one definition, so both produce the same FP32 features).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.hpp"
#include "kernels.hpp"

namespace nwprt {
namespace dev {

// Spatial hash of Instant-NGP (Mueller et al. 2022): x*1 ^ y*2654435761 ^
// z*805459861 in uint32 arithmetic.
__device__ __forceinline__ uint32_t grid_hash(uint32_t x, uint32_t y, uint32_t z) {
  return x ^ (y * 2654435761u) ^ (z * 805459861u);
}

// Level l at position x: the 8 corner entry indices within the level and
// the trilinear fractions.  Position pos = fmaf(x, scale_l, 0.5f) in FP32;
// corner c takes bit 0 / 1 / 2 for +x / +y / +z; dense levels index
// x + y res + z res^2 (mod size), hashed ones grid_hash & (size - 1).
__device__ __forceinline__ void level_corners(const nwprt_hashgrid_levels &lv, int l,
                                              const float (&x)[3], uint32_t (&idx)[8],
                                              float (&fr)[3]) {
  const float scale = lv.scale[l];
  const uint32_t res = lv.resolution[l], size = lv.size[l];
  const bool dense = lv.dense[l] != 0;
  uint32_t gi[3];
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
    uint32_t i;
    if (dense) {
      i = cx + cy * res + cz * res * res;
      if (i >= size) i %= size;
    } else {
      i = grid_hash(cx, cy, cz) & (size - 1u);
    }
    idx[c] = i;
  }
}

// Trilinear weight of corner c.
__device__ __forceinline__ float corner_weight(int c, const float (&fr)[3]) {
  return ((c & 1) ? fr[0] : 1.0f - fr[0]) * ((c & 2) ? fr[1] : 1.0f - fr[1]) *
         ((c & 4) ? fr[2] : 1.0f - fr[2]);
}

// Two features of one corner entry (F = 2): FP16 pair or FP32 pair.
template <bool HF16>
struct Entry2 {
  typedef typename std::conditional<HF16, uint32_t, float2>::type raw;
  __device__ static __forceinline__ raw load(const void *hash, size_t e) {
    if constexpr (HF16)
      return *((gptr<uint32_t>)hash + e);
    else
      return __builtin_bit_cast(raw, *((gptr<unsigned long long>)hash + e));
  }
  __device__ static __forceinline__ void add(float w, raw v, float *acc) {
    if constexpr (HF16) {
      acc[0] = fmaf(w, dev::f16_to_f32((uint16_t)(v & 0xffffu)), acc[0]);
      acc[1] = fmaf(w, dev::f16_to_f32((uint16_t)(v >> 16)), acc[1]);
    } else {
      acc[0] = fmaf(w, v.x, acc[0]);
      acc[1] = fmaf(w, v.y, acc[1]);
    }
  }
};

// Trilinear hash-grid features of level l at position x, added to acc[0..F)
// corner by corner (c = 0..7, one FMA per feature).  idx (may be null)
// receives the 8 indices.
template <int F, bool HF16>
__device__ __forceinline__ void accumulate_level(const nwprt_hashgrid_levels &lv, const void *hash,
                                                 int l, const float (&x)[3], float *acc,
                                                 uint32_t *idx_out) {
  uint32_t idx[8];
  float fr[3];
  level_corners(lv, l, x, idx, fr);
  const uint32_t off = lv.offset[l];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float w = corner_weight(c, fr);
    if (F == 2) {
      Entry2<HF16>::add(w, Entry2<HF16>::load(hash, (size_t)off + idx[c]), acc);
    } else {
      const size_t e = ((size_t)off + idx[c]) * F;
#pragma unroll
      for (int j = 0; j < F; ++j)
        acc[j] = fmaf(w, HF16 ? dev::f16_to_f32(((const uint16_t *)hash)[e + j])
                              : ((const float *)hash)[e + j],
                      acc[j]);
    }
    if (idx_out) idx_out[c] = idx[c];
  }
}

// Levels l0 .. l0 + NL - 1 (F = 2) at x, each from zero: all NL x 8 corner
// loads are issued before the first is used, so a wave waits for memory
// once per batch of levels, not once per level.  The per-level arithmetic
// is accumulate_level's, in the same order.
template <int NL, bool HF16>
__device__ __forceinline__ void gather_levels(const nwprt_hashgrid_levels &lv, const void *hash,
                                              int l0, const float (&x)[3], float (&out)[NL][2]) {
  typename Entry2<HF16>::raw v[NL][8];
  float fr[NL][3];
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    uint32_t idx[8];
    level_corners(lv, l0 + i, x, idx, fr[i]);
    const uint32_t off = lv.offset[l0 + i];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[i][c] = Entry2<HF16>::load(hash, (size_t)off + idx[c]);
  }
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    out[i][0] = 0.f;
    out[i][1] = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) Entry2<HF16>::add(corner_weight(c, fr[i]), v[i][c], out[i]);
  }
}

}  // namespace dev
}  // namespace nwprt
