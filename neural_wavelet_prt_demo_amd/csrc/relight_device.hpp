// Device pieces shared by the relight kernels (BF16/FP16 and FP8 forms).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.hpp"
#include "hashgrid_device.hpp"
#include "kernels.hpp"

namespace nwprt {
namespace dev {

// Per-pixel outputs of computeResponse (Renderer/tinycudann/utils_kernels.h:80-88):
// indirect = (rgb, 1), indirect + direct = (rgb + direct, 1), and the RGBA8
// tone map of the chosen one (Renderer/toneMap.cu:47-71).
__device__ __forceinline__ void write_outputs(const RelightArgs &a, int p, float r, float g,
                                              float b) {
  const float4 d = *(const float4 *)(a.direct + 4 * (size_t)p);
  const float4 ind = make_float4(r, g, b, 1.0f);
  const float4 wd = make_float4(r + d.x, g + d.y, b + d.z, 1.0f);
  if (a.indirect) *(float4 *)(a.indirect + 4 * (size_t)p) = ind;
  if (a.with_direct) *(float4 *)(a.with_direct + 4 * (size_t)p) = wd;
  if (a.rgba8 && a.tonemap_source != NWPRT_TONEMAP_NONE) {
    const float4 t = a.tonemap_source == NWPRT_TONEMAP_FULL ? wd : ind;
    a.rgba8[p] = tonemap_rgba8(t.x, t.y, t.z);
  }
}

// Eight 16-bit features of pixel p at contraction slots slot0 .. slot0 + 7
// (slot0 a multiple of 8): one 16-byte row read, or for planar features two
// 8-byte reads from planes slot0 / 4 and slot0 / 4 + 1 (plane stride P).
__device__ __forceinline__ u32x4 load_features8(const RelightArgs &a, int P, int p, int slot0) {
  if (a.features_planar) {
    const gptr<uint64_t> pl = (gptr<uint64_t>)a.features;
    const uint64_t lo = pl[(size_t)(slot0 >> 2) * P + p];
    const uint64_t hi = pl[(size_t)((slot0 >> 2) + 1) * P + p];
    return u32x4{(uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32)};
  }
  return *(gptr<u32x4>)((const uint16_t *)a.features + (size_t)p * 32 + slot0);
}

// Transport of one output (softplus or identity); its clamped light term is
// max(light * T, 0) (Renderer/tinycudann/utils_kernels.h:74-78).
template <bool SOFTPLUS>
__device__ __forceinline__ float transport(float o) {
  return SOFTPLUS ? softplus(o) : o;
}

// Clamped light terms max(light * T, 0) of the MFMA forms, accumulated with
// two VALU operations fewer per term than the literal form:
// - softplus: T = ln 2 * log2(1 + 2^y), y = o log2(e), and T >= 0, so the
//   term is max(light, 0) * T.  The clamp is applied to the wave-uniform
//   light (a scalar operation) and the factor ln 2 to the pixel's final sum
//   (`kSumScale`) instead of every term; above o = 15 the scaled transport
//   is y itself (1 + 2^y rounds to 2^y there);
// - identity: T = o may be negative and the term stays max(light * o, 0).
// Every term is non-negative either way.
template <bool SOFTPLUS>
struct LightTerm {
  static constexpr float kSumScale = SOFTPLUS ? 0.693147180559945309f : 1.0f;
  // Per wavelet and channel: the light as the term uses it.
  __device__ static __forceinline__ float light(float l) { return SOFTPLUS ? relu_f32(l) : l; }
  // SOFTPLUS: the transport divided by kSumScale, the factor the light
  // multiplies in term().
  __device__ static __forceinline__ float factor(float o) {
    const float y = o * 1.44269504088896341f;
    const float sp = __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(y));
    return o > 15.0f ? y : sp;
  }
  // Per pair: the term divided by kSumScale.
  __device__ static __forceinline__ float term(float light, float o) {
    if constexpr (SOFTPLUS) {
      const float y = o * 1.44269504088896341f;
      const float sp = __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(y));
      return light * (o > 15.0f ? y : sp);
    } else {
      return relu_f32(light * o);
    }
  }
};

// Kahan-compensated accumulation of non-negative terms: s - c is the
// running sum (c holds the low-order part lost by the last addition, with
// its sign flipped).  Contraction is off: a product that produced x is
// never fused into x - c, whatever the caller's code shape, so kernel forms
// that order their work differently sum identically.
__device__ __forceinline__ void kahan_add(float &s, float &c, float x) {
#pragma clang fp contract(off)
  const float y = x - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// Scheduling hint for one pipeline stage: NM MFMAs, each followed by its
// share of the stage's NV vector-ALU instructions, so the matrix core and
// the VALU work on different pixel groups at the same time.
template <int NM, int NV>
__device__ __forceinline__ void interleave_stage() {
#pragma unroll
  for (int i = 0; i < NM; ++i) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                  // MFMA
    __builtin_amdgcn_sched_group_barrier(0x002, (NV + NM - 1) / NM, 0);  // VALU
  }
  __builtin_amdgcn_sched_barrier(0);
}

// Levels l0 .. l0 + NL - 1 of one pixel, packed into the planar slot order
// (u32 2q = level q, 2q + 1 = level 15 - q).
template <typename E, int NL, bool HF16>
__device__ __forceinline__ void encode_batch(const RelightArgs &a, int l0, const float (&x)[3],
                                             uint32_t (&pk)[16]) {
  float v[NL][2];
  gather_levels<NL, HF16>(a.lv, a.hash, l0, x, v);
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const int l = l0 + i;
    pk[l < 8 ? 2 * l : 2 * (15 - l) + 1] = std::is_same<E, _Float16>::value
                                                ? pack_f16_rounded(v[i][0], v[i][1])
                                                : pack_bf16(v[i][0], v[i][1]);
  }
}

// Fused encode of a wave's 64 pixels: lane l computes the 16 levels of
// pixel my_px (level parameters uniform across the wave, as in the
// standalone encode, with the same lookup code), packs each level's two
// features in the planar slot order and stores the 16 words to its 64 B row
// of `lds` ([64 pixels][16 words]).  Inactive lanes (past the frame or not
// hit) store zeros.  On return every lane may read any row: the stores are
// ordered before the caller's loads by a wavefront fence (no workgroup
// barrier, so waves that exit early are harmless).
template <typename E>
__device__ __forceinline__ void fused_encode_lds(const RelightArgs &a, bool active, int my_px,
                                                 int lane, uint32_t *lds) {
  uint32_t pk[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) pk[i] = 0u;
  if (active) {
    const f32x4 x4 = *(gptr<f32x4>)(a.position + 4 * (size_t)my_px);
    const float x[3] = {x4[0], x4[1], x4[2]};
    // Four levels per batch: 32 gathers in flight before the first is used
    // (FP16 entries; two levels of FP32 entries per half-batch).  The table
    // type is tested per batch: that branch also keeps the compiler from
    // hoisting the address math of all 16 levels at once (which spills).
#pragma unroll
    for (int l0 = 0; l0 < 16; l0 += 4) {
      if (a.hash_f16) {
        encode_batch<E, 4, true>(a, l0, x, pk);
      } else {
        encode_batch<E, 2, false>(a, l0, x, pk);
        encode_batch<E, 2, false>(a, l0 + 2, x, pk);
      }
    }
  }
  u32x4 *row = (u32x4 *)(lds + lane * 16);
#pragma unroll
  for (int c = 0; c < 4; ++c) row[c] = u32x4{pk[4 * c], pk[4 * c + 1], pk[4 * c + 2], pk[4 * c + 3]};
  // Lanes of one wave exchange through LDS: order the stores before the loads.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace dev
}  // namespace nwprt
