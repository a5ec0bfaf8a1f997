// Wave-level pieces of the BF16 / FP16 fused relight (v_mfma_f32_16x16x32):
// the MFMA wrappers, ReLU + pack, one wavelet's four layers (plain and
// software-pipelined), the fused encode's operand read and the per-wave
// relight.  Shared by the BF16 / FP16 one-call, batch and allocation-free
// kernels and by the FP8 form's calibration pass.
//
// MFMA data layout (v_mfma_f32_16x16x32_{bf16,f16}; lane l, g = l/16,
// n = l%16): the A operand holds row n, k-slots 8g..8g+7; B holds column n,
// the same k-slots; the result D holds rows 4g..4g+3 of column n.  Weights
// are the A operand (rows = output features) and pixels are the columns, so
// a layer's D block (output features 16b + 4g + j) is already laid out as
// the next layer's B operand up to a permutation of the contraction index,
// which the host folds into the weight tiles (build_weight_tiles).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.hpp"
#include "hashgrid_device.hpp"
#include "kernels.hpp"
#include "relight_device.hpp"

namespace nwprt {
namespace mma16 {

using dev::f32x4;
using dev::u32x4;

template <typename E>
struct Mfma;
template <>
struct Mfma<__bf16> {
  static __device__ __forceinline__ f32x4 run(u32x4 a, u32x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(dev::bf16x8, a),
                                                   __builtin_bit_cast(dev::bf16x8, b), c, 0, 0, 0);
  }
  static __device__ __forceinline__ uint32_t pack(float lo, float hi) {
    return dev::pack_bf16(lo, hi);
  }
};
template <>
struct Mfma<_Float16> {
  static __device__ __forceinline__ f32x4 run(u32x4 a, u32x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(dev::f16x8, a),
                                                  __builtin_bit_cast(dev::f16x8, b), c, 0, 0, 0);
  }
  static __device__ __forceinline__ uint32_t pack(float lo, float hi) {
    return dev::pack_f16(lo, hi);
  }
};

// 16x16x16 MFMA for the 16-wide aux block of the second layer: lane (g, n)
// holds k-slots 4g..4g+3 of row / column n.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
template <typename E>
struct Mfma16;
template <>
struct Mfma16<__bf16> {
  static __device__ __forceinline__ f32x4 run(u32x2 a, u32x2 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(bf16x4, a),
                                                     __builtin_bit_cast(bf16x4, b), c, 0, 0, 0);
  }
};
template <>
struct Mfma16<_Float16> {
  static __device__ __forceinline__ f32x4 run(u32x2 a, u32x2 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(f16x4, a),
                                                 __builtin_bit_cast(f16x4, b), c, 0, 0, 0);
  }
};

// ReLU + round of two 16-row D blocks (output features 32kb + [0,32)) into
// one B operand: slots 0..3 <- block 2kb rows 4g..4g+3, slots 4..7 <- block
// 2kb+1 rows 4g..4g+3.
template <typename E>
__device__ __forceinline__ u32x4 relu_pack(f32x4 lo, f32x4 hi) {
  u32x4 r;
  r[0] = dev::cvt_relu2<E>(lo[0], lo[1]);
  r[1] = dev::cvt_relu2<E>(lo[2], lo[3]);
  r[2] = dev::cvt_relu2<E>(hi[0], hi[1]);
  r[3] = dev::cvt_relu2<E>(hi[2], hi[3]);
  return r;
}

using dev::transport;
using dev::write_outputs;

// The four layers for one selected wavelet (A operands `a1` = its
// first-layer tiles) over the wave's 64 pixels.  Returns the output layer's
// accumulator: lane l holds the three pre-activation outputs of pixel
// base + l in elements 0..2.
template <typename E, bool AUX, int GP, bool W4LDS>
__device__ __forceinline__ f32x4 wavelet_step(const u32x4 (&a1)[4], const u32x4 (&hb)[4],
                                              const u32x2 (&auxb)[4], const u32x2 (&waux)[4],
                                              const u32x4 (&w2)[4][2], const u32x4 (&w3)[4][2],
                                              const u32x4 (&w4)[4][2], const u32x4 *w4s,
                                              int lane) {
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  f32x4 out = zero;
#pragma unroll
  for (int jp = 0; jp < 4; jp += GP) {
    f32x4 c[GP][4];
    u32x4 x[GP][2];
#pragma unroll
    for (int j = 0; j < GP; ++j)
#pragma unroll
      for (int b = 0; b < 4; ++b) c[j][b] = Mfma<E>::run(a1[b], hb[jp + j], zero);
#pragma unroll
    for (int j = 0; j < GP; ++j) {
      x[j][0] = relu_pack<E>(c[j][0], c[j][1]);
      x[j][1] = relu_pack<E>(c[j][2], c[j][3]);
    }
#pragma unroll
    for (int j = 0; j < GP; ++j)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        // Second layer; with aux inputs its 16 extra columns enter last, as
        // one more link of the same accumulation chain.
        c[j][b] = Mfma<E>::run(w2[b][1], x[j][1], Mfma<E>::run(w2[b][0], x[j][0], zero));
        if (AUX) c[j][b] = Mfma16<E>::run(waux[b], auxb[jp + j], c[j][b]);
      }
#pragma unroll
    for (int j = 0; j < GP; ++j) {
      x[j][0] = relu_pack<E>(c[j][0], c[j][1]);
      x[j][1] = relu_pack<E>(c[j][2], c[j][3]);
    }
#pragma unroll
    for (int j = 0; j < GP; ++j)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        c[j][b] = Mfma<E>::run(w3[b][1], x[j][1], Mfma<E>::run(w3[b][0], x[j][0], zero));
#pragma unroll
    for (int j = 0; j < GP; ++j) {
      x[j][0] = relu_pack<E>(c[j][0], c[j][1]);
      x[j][1] = relu_pack<E>(c[j][2], c[j][3]);
    }
    // Output layer: variant j carries W4 in rows 4j..4j+2, so group j's
    // three outputs land in rows 4j.. of the shared accumulator, i.e. in
    // lanes with g == j -- the lanes that own those pixels.
#pragma unroll
    for (int j = 0; j < GP; ++j) {
      const u32x4 t0 = W4LDS ? w4s[((jp + j) * 2 + 0) * 64 + lane] : w4[jp + j][0];
      const u32x4 t1 = W4LDS ? w4s[((jp + j) * 2 + 1) * 64 + lane] : w4[jp + j][1];
      out = Mfma<E>::run(t1, x[j][1], Mfma<E>::run(t0, x[j][0], out));
    }
  }
  return out;
}

// Software-pipelined form of wavelet_step: the four pixel groups flow
// through the four layers two at a time, and every stage pairs one group's
// MFMAs (M) with another group's ReLU + pack (C):
//   M1(0) | M1(1) C(0) | M2(0) C(1) | M2(1) C(0) | M3(0) C(1) | M3(1) C(0) |
//   M4(0) M1(2) C(1) | M4(1) M1(3) C(2) | M2(2) C(3) | ... | M4(3)
// Per-lane arithmetic and its order are those of wavelet_step.
template <typename E, bool AUX, bool W4LDS>
__device__ __forceinline__ f32x4 wavelet_step_pipelined(
    const u32x4 (&a1)[4], const u32x4 (&hb)[4], const u32x2 (&auxb)[4], const u32x2 (&waux)[4],
    const u32x4 (&w2)[4][2], const u32x4 (&w3)[4][2], const u32x4 (&w4)[4][2], const u32x4 *w4s,
    int lane) {
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  f32x4 out = zero;
  f32x4 c[4][4];
  u32x4 x[4][2];
  constexpr int kAux = AUX ? 4 : 0;
  auto M1 = [&](int g) {
#pragma unroll
    for (int b = 0; b < 4; ++b) c[g][b] = Mfma<E>::run(a1[b], hb[g], zero);
  };
  // Layers 2 and 3: the four first-half-K MFMAs are issued before the four
  // second-half ones, so no MFMA waits on the one just issued.
  auto M2 = [&](int g) {
    f32x4 t[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) t[b] = Mfma<E>::run(w2[b][0], x[g][0], zero);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      c[g][b] = Mfma<E>::run(w2[b][1], x[g][1], t[b]);
      if (AUX) c[g][b] = Mfma16<E>::run(waux[b], auxb[g], c[g][b]);
    }
  };
  auto M3 = [&](int g) {
    f32x4 t[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) t[b] = Mfma<E>::run(w3[b][0], x[g][0], zero);
#pragma unroll
    for (int b = 0; b < 4; ++b) c[g][b] = Mfma<E>::run(w3[b][1], x[g][1], t[b]);
  };
  auto M4 = [&](int g) {
    const u32x4 t0 = W4LDS ? w4s[(g * 2 + 0) * 64 + lane] : w4[g][0];
    const u32x4 t1 = W4LDS ? w4s[(g * 2 + 1) * 64 + lane] : w4[g][1];
    out = Mfma<E>::run(t1, x[g][1], Mfma<E>::run(t0, x[g][0], out));
  };
  auto C = [&](int g) {
    x[g][0] = relu_pack<E>(c[g][0], c[g][1]);
    x[g][1] = relu_pack<E>(c[g][2], c[g][3]);
  };
  M1(0);
  __builtin_amdgcn_sched_barrier(0);
  M1(1); C(0); dev::interleave_stage<4, 16>();
  M2(0); C(1); dev::interleave_stage<8 + kAux, 16>();
  M2(1); C(0); dev::interleave_stage<8 + kAux, 16>();
  M3(0); C(1); dev::interleave_stage<8, 16>();
  M3(1); C(0); dev::interleave_stage<8, 16>();
  M4(0); M1(2); C(1); dev::interleave_stage<6, 16>();
  M4(1); M1(3); C(2); dev::interleave_stage<6, 16>();
  M2(2); C(3); dev::interleave_stage<8 + kAux, 16>();
  M2(3); C(2); dev::interleave_stage<8 + kAux, 16>();
  M3(2); C(3); dev::interleave_stage<8, 16>();
  M3(3); C(2); dev::interleave_stage<8, 16>();
  M4(2); C(3); dev::interleave_stage<2, 16>();
  M4(3);
  return out;
}

// Fused encode of a wave's 64 pixels into the planar B operands: each lane
// encodes its own pixel into the wave's LDS rows (dev::fused_encode_lds);
// lane (g, n) then reads slots 4g .. 4g + 3 of pixel 16 j + n for group j --
// the words load_features8 reads from the planar feature planes.  Pixels
// that are past the frame or not hit get zero features.
template <typename E>
__device__ __forceinline__ void fused_features(const RelightArgs &a, bool active, int my_px,
                                               int lane, int g, int n, uint32_t *lds,
                                               u32x4 (&hb)[4]) {
  dev::fused_encode_lds<E>(a, active, my_px, lane, lds);
#pragma unroll
  for (int j = 0; j < 4; ++j) hb[j] = *(const u32x4 *)(lds + (16 * j + n) * 16 + 4 * g);
}

// One wave's 64 pixels [base, base + 64) of a frame (base < a.n_pixels).
// Pixel groups j = 0..3 (16 pixels each) are the MFMA columns; lane l owns
// pixel base + l for the reduction.  The next wavelet's first-layer tiles
// are fetched while the current one computes (two register sets, K loop
// unrolled by two).  GP = 0 selects the software-pipelined step; W4LDS reads
// the output layer's tiles from `w4s` (filled by the caller).
// FUSE: the fused-encode path is compiled in (taken when a.fused_encode).
template <typename E, bool SOFTPLUS, bool DUMP, bool AUX, int GP, bool W4LDS, bool FUSE>
__device__ __forceinline__ void relight_wave(const RelightArgs &a, int base, int lane,
                                             const u32x4 *w4s, uint32_t *feat_lds) {
  const int g = lane >> 4, n = lane & 15;
  const dev::gptr<u32x4> wt = (dev::gptr<u32x4>)a.w_tiles;
  const int P = a.n_pixels;
  const int my_px = base + lane;
  const bool mine = my_px < P;
  const bool hit =
      mine && (a.env_flag == nullptr || *((dev::gptr<uint8_t>)a.env_flag + my_px) != 0);
  const bool any_hit = __ballot(hit) != 0;

  float sum[3] = {0.f, 0.f, 0.f}, comp[3] = {0.f, 0.f, 0.f};
  if (any_hit) {
    // Per-pixel features as B operands: group j, column n, k-slots 8g..8g+7.
    u32x4 hb[4];
    if (FUSE && a.fused_encode) {
      fused_features<E>(a, mine && hit, my_px, lane, g, n, feat_lds, hb);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int p = base + 16 * j + n;
        hb[j] = p < P ? dev::load_features8(a, P, p, 8 * g) : u32x4{0u, 0u, 0u, 0u};
      }
    }
    // Aux inputs (B: group j, column n, k-slots 4g..4g+3) and the second
    // layer's aux columns (A tiles [4][64]).
    u32x2 auxb[4], waux[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int p = base + 16 * j + n;
      auxb[j] = AUX && p < P ? *(const u32x2 *)((const uint16_t *)a.aux + (size_t)p * 16 + 4 * g)
                             : u32x2{0u, 0u};
      waux[j] = AUX ? ((const u32x2 *)a.waux_tiles)[j * 64 + lane] : u32x2{0u, 0u};
    }
    // Layers 2..4 as A-operand tiles: [4][2][64] each.
    u32x4 w2[4][2], w3[4][2], w4[4][2];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
        w2[b][kb] = wt[(0 * 8 + b * 2 + kb) * 64 + lane];
        w3[b][kb] = wt[(1 * 8 + b * 2 + kb) * 64 + lane];
        if (!W4LDS) w4[b][kb] = wt[(2 * 8 + b * 2 + kb) * 64 + lane];
      }
    const dev::gptr<u32x4> mk = (dev::gptr<u32x4>)a.mk;
    u32x4 ta[4], tb[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) ta[b] = mk[b * 64 + lane];
    auto step = [&](const u32x4(&t)[4]) -> f32x4 {
      if constexpr (GP == 0)
        return wavelet_step_pipelined<E, AUX, W4LDS>(t, hb, auxb, waux, w2, w3, w4, w4s, lane);
      else
        return wavelet_step<E, AUX, GP, W4LDS>(t, hb, auxb, waux, w2, w3, w4, w4s, lane);
    };
    // Wavelets are taken two at a time: the clamped terms of k and k + 1 are
    // added, and that pair sum enters the Kahan-compensated running sum.
    using LT = dev::LightTerm<SOFTPLUS>;
    for (int k = 0; k < a.k; k += 2) {
      const bool two = k + 1 < a.k;
      if (two) {
#pragma unroll
        for (int b = 0; b < 4; ++b) tb[b] = mk[((size_t)(k + 1) * 4 + b) * 64 + lane];
      }
      const f32x4 o0 = step(ta);
      f32x4 o1 = {0.f, 0.f, 0.f, 0.f};
      if (two) {
        if (k + 2 < a.k) {
#pragma unroll
          for (int b = 0; b < 4; ++b) ta[b] = mk[((size_t)(k + 2) * 4 + b) * 64 + lane];
        }
        o1 = step(tb);
      }
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        float x = LT::term(LT::light(dev::load_invariant(a.light + 3 * k + ch)), o0[ch]);
        if (two) x += LT::term(LT::light(dev::load_invariant(a.light + 3 * (k + 1) + ch)), o1[ch]);
        dev::kahan_add(sum[ch], comp[ch], x);
        if (DUMP) {
          const int q = my_px - a.dump_begin;
          if (mine && q >= 0 && q < a.dump_count) {
            a.dump[((size_t)q * a.k + k) * 3 + ch] = transport<SOFTPLUS>(o0[ch]);
            if (two) a.dump[((size_t)q * a.k + k + 1) * 3 + ch] = transport<SOFTPLUS>(o1[ch]);
          }
        }
      }
    }
  }
  if (mine) {
    constexpr float f = dev::LightTerm<SOFTPLUS>::kSumScale;
    const float r = hit ? (sum[0] - comp[0]) * f : 0.f, gg = hit ? (sum[1] - comp[1]) * f : 0.f,
                bb = hit ? (sum[2] - comp[2]) * f : 0.f;
    write_outputs(a, my_px, r, gg, bb);
  }
}

}  // namespace mma16
}  // namespace nwprt
