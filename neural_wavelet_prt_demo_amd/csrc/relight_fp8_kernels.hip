// FP8 form of the fused per-(pixel, wavelet) network for gfx950.
//
// Same computation as the BF16 kernel (Renderer/tinycudann/tcnn.cu:111-152,
// utils_kernels.h:51-115 replaced by one launch), with the hidden layers on
// v_mfma_scale_f32_32x32x64_f8f6f4: OCP E4M3 operands, FP32 accumulation,
// one power-of-two scale per layer applied by the instruction's E8M0 scale
// operands.  The first layer (the per-wavelet Hadamard product folded into
// W1) stays BF16 on v_mfma_f32_32x32x16_bf16: its contraction is only 32
// long, so an FP8 K=64 instruction would be half padding.
//
// 32x32 MFMA layouts (gfx950, lane l, h = l / 32, c = l % 32):
//   A: row c, contraction slots (h, j) for its bytes / elements j;
//   B: column c, the same slots;
//   D: column c, element v holds row 8 (v / 4) + 4 h + v % 4.
// A hidden layer's two 32-row D blocks hold, in lane l, 32 of the 64 outputs
// of pixel column c: exactly the 32 slots B needs from that lane.  Slot
// (h, j) of the next layer therefore carries output
//   pi(h, j) = 32 (j / 16) + 8 ((j % 16) / 4) + 4 h + j % 4,
// and the host permutes the weight columns to match (build_fp8_weight_tiles).
//
// Scaling.  Weights: W * 2^ew quantized, ew per layer from the weight amax.
// Activations: the maxima of the BF16 network's layer 1..3 outputs over a
// pixel subsample of about 16K pixels (calib_amax_kernel) give ea per layer,
// with amax * 2^ea in (7/64, 7/32].  Layer 1's 2^ea1 is applied to the
// per-pixel features once; layer n's MFMA scales (A: 2^(ea_n - ew_n),
// B: 2^-(ea_(n-1) + 8)) leave its output already multiplied by 2^ea_n.  The
// ReLU + convert step is then two instructions per pair of outputs:
// v_cvt_pk_f16_f32 with its clamp to [0, 1] (ReLU and saturation at once;
// the compiler folds the clamp into the conversion) and
// v_cvt_scalef32_pk_fp8_f16, which rounds v * 2^8 to E4M3.  The E4M3
// operand of layer n + 1 is therefore the activation times 2^(ea_n + 8):
// the sampled maximum lands in (28, 56], activations up to 4.5x it convert
// exactly, anything beyond 9.1x it saturates at 256 (never NaN), and only
// values under 2^-6 -- below 5.6e-4 of the sampled maximum -- lose bits to
// the reserve as E4M3 subnormals.  The FP16 step rounds to 11 bits first,
// so the E4M3 byte differs from a direct rounding only where the FP16 value
// falls exactly on an E4M3 rounding tie.  Every scale is a power of two:
// multiplying the inputs by 2^s multiplies the result by exactly 2^s.
#include <hip/hip_runtime.h>

#include <cmath>

#include "device_common.hpp"
#include "hashgrid_device.hpp"
#include "kernels.hpp"
#include "relight_device.hpp"

namespace nwprt {
namespace {

using dev::u32x4;
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef short i16x2 __attribute__((ext_vector_type(2)));

constexpr int kBlock = 256;  // 4 waves x 64 pixels

__device__ __forceinline__ f32x16 mfma_bf16(u32x4 a, u32x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(dev::bf16x8, a),
                                                 __builtin_bit_cast(dev::bf16x8, b), c, 0, 0, 0);
}

// E4M3 x E4M3 (format code 0 for both operands), A scaled by 2^(sa - 127),
// B by 2^(sb - 127).
__device__ __forceinline__ f32x16 mfma_fp8(i32x8 a, i32x8 b, f32x16 c, int sa, int sb) {
  return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, sa, 0, sb);
}

// ReLU, saturation and E4M3 rounding of a hidden layer's two D blocks into
// the next layer's B operand: dword d holds block d / 4, elements
// 4 (d % 4) .. + 3, i.e. slots j = 4 d .. 4 d + 3.  Each pair is rounded to
// FP16 and clamped to [0, 1] by one v_cvt_pk_f16_f32 (the min / max below
// fold into its clamp bit), then converted to E4M3 times 2^8 by one
// v_cvt_scalef32_pk_fp8_f16 (which divides by its scale operand, 2^-8).
__device__ __forceinline__ uint32_t clamp01_f16x2(float a, float b) {
  const dev::f32x2 v = {a, b};
  const dev::f16x2 h = __builtin_convertvector(v, dev::f16x2);
  const dev::f16x2 z = {(_Float16)0.f, (_Float16)0.f}, o = {(_Float16)1.f, (_Float16)1.f};
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_elementwise_max(h, z), o));
}

__device__ __forceinline__ i32x8 relu_fp8(const f32x16 &lo, const f32x16 &hi) {
  i32x8 x;
#pragma unroll
  for (int d = 0; d < 8; ++d) {
    const f32x16 &src = d < 4 ? lo : hi;
    const int v = 4 * (d & 3);
    const uint32_t h0 = clamp01_f16x2(src[v + 0], src[v + 1]);
    const uint32_t h1 = clamp01_f16x2(src[v + 2], src[v + 3]);
    // The first convert fills the low word; its high word is replaced by the
    // second, so its "old" operand may be anything: h0's own bits let the
    // result take h0's register instead of a zeroed one (no move).
    i16x2 w = __builtin_amdgcn_cvt_scalef32_pk_fp8_f16(
        __builtin_bit_cast(i16x2, h0), __builtin_bit_cast(dev::f16x2, h0), 0x1p-8f, false);
    w = __builtin_amdgcn_cvt_scalef32_pk_fp8_f16(w, __builtin_bit_cast(dev::f16x2, h1), 0x1p-8f,
                                                 true);
    x[d] = __builtin_bit_cast(int, w);
  }
  return x;
}

// Exponent e with amax * 2^e in (7/64, 7/32], from the float's bits (so it
// moves by exactly -s when amax is multiplied by 2^s).  0 for a zero,
// denormal or non-finite maximum; clamped to [-100, 100].
__device__ __forceinline__ int act_exp(float amax) {
  const uint32_t u = __float_as_uint(amax);
  const int be = (int)((u >> 23) & 0xffu);
  if (be == 0 || be == 255 || (u >> 31)) return 0;
  const int E = be - 127;
  const int e = (u & 0x7fffffu) <= 0x600000u ? -3 - E : -4 - E;
  return min(max(e, -100), 100);
}

__device__ __forceinline__ int scale_byte(int e) { return min(max(e, 0), 254); }

// Eight BF16 features times 2^e (exact: a power of two), re-rounded to BF16.
__device__ __forceinline__ u32x4 scale_bf16x8(u32x4 v, float s) {
  u32x4 r;
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    const float lo = __uint_as_float(v[d] << 16) * s;
    const float hi = __uint_as_float(v[d] & 0xffff0000u) * s;
    r[d] = dev::pack_bf16(lo, hi);
  }
  return r;
}

// One wave's 64 pixels [base, base + 64) of a frame (base < a.n_pixels), as
// two 32-column groups; after the output layer lane l holds pixel base + l.
// The next wavelet's first-layer tiles load while the current one computes.
// `w4s`: the wave's own copy of the output layer's two tiles (lane l reads
// only entries l and 64 + l, which it stored itself).  AUX: the second
// layer's 16 aux columns enter as a BF16 32x32x16 MFMA accumulated onto the
// FP8 layer's result (the last link of its chain); its B operand, the
// pixel's aux inputs, is pre-multiplied by 2^ea2, so it adds in the scaled
// domain of that result.  FUSE: the fused-encode path is compiled in (taken
// when a.fused_encode).
// term(k) + term(k + 1) of one channel, the second product fused into the
// addition (one rounding), as every FP8 variant forms the pair.
template <bool SOFTPLUS>
__device__ __forceinline__ float pair_sum(float first, float light, float o) {
  using LT = dev::LightTerm<SOFTPLUS>;
  if constexpr (SOFTPLUS) return fmaf(light, LT::factor(o), first);
  else return first + LT::term(light, o);
}

// Wave forms (MODE): 0 one wavelet's four layers for both pixel groups at
// a time, two wavelets per iteration; 1 the same with the two groups
// software-pipelined; 2 ("seq", for three waves per SIMD) one wavelet per
// iteration, the groups taken through layers 2-4 one after the other, and
// the next wavelet's first-layer tiles loaded as soon as layer 1 is done.
// Every form computes the same per-lane values in the same order.
template <bool SOFTPLUS, bool DUMP, bool AUX, bool FUSE, int MODE>
__device__ __forceinline__ void relight_fp8_wave(const RelightArgs &a, int base, int lane,
                                                 const i32x8 *w4s, uint32_t *feat_lds) {
  const int h = lane >> 5, col = lane & 31;
  const int P = a.n_pixels;
  const int my_px = base + lane;
  const bool mine = my_px < P;
  const bool hit =
      mine && (a.env_flag == nullptr || *((dev::gptr<uint8_t>)a.env_flag + my_px) != 0);
  const bool any_hit = __ballot(hit) != 0;

  float sum[3] = {0.f, 0.f, 0.f}, comp[3] = {0.f, 0.f, 0.f};
  if (any_hit) {
    const int ea1 = act_exp(a.amax[0]), ea2 = act_exp(a.amax[1]), ea3 = act_exp(a.amax[2]);
    // B operands carry the previous layer's activations times 2^(ea + 8).
    const int sa2 = scale_byte(127 + ea2 - a.w8_exp[0]), sb2 = scale_byte(119 - ea1);
    const int sa3 = scale_byte(127 + ea3 - a.w8_exp[1]), sb3 = scale_byte(119 - ea2);
    const int sa4 = scale_byte(127 - a.w8_exp[2]), sb4 = scale_byte(119 - ea3);
    const float fs = __uint_as_float((uint32_t)(ea1 + 127) << 23);

    // First-layer B operands: group g, column col, slots 16 s + 8 h + [0, 8).
    u32x4 f[2][2];
    if (FUSE && a.fused_encode) {
      dev::fused_encode_lds<__bf16>(a, hit, my_px, lane, feat_lds);
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int s = 0; s < 2; ++s)
          f[g][s] = scale_bf16x8(*(const u32x4 *)(feat_lds + (32 * g + col) * 16 + 8 * s + 4 * h),
                                 fs);
    } else {
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        const int p = base + 32 * g + col;
#pragma unroll
        for (int s = 0; s < 2; ++s)
          f[g][s] = p < P ? scale_bf16x8(dev::load_features8(a, P, p, 16 * s + 8 * h), fs)
                          : u32x4{0u, 0u, 0u, 0u};
      }
    }
    // Aux inputs (B: group g, column col, inputs 8 h .. 8 h + 7, times
    // 2^ea2) and the second layer's aux columns (A tiles [2][64]).
    u32x4 auxb[2], waux[2];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int p = base + 32 * g + col;
      auxb[g] = AUX && p < P
                    ? scale_bf16x8(*(dev::gptr<u32x4>)((const uint16_t *)a.aux + (size_t)p * 16 + 8 * h),
                                   __uint_as_float((uint32_t)(ea2 + 127) << 23))
                    : u32x4{0u, 0u, 0u, 0u};
      waux[g] = AUX ? ((dev::gptr<u32x4>)a.waux32_tiles)[g * 64 + lane] : u32x4{0u, 0u, 0u, 0u};
    }
    const dev::gptr<i32x8> w8 = (dev::gptr<i32x8>)a.w8_tiles;
    i32x8 w2[2], w3[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      w2[b] = w8[(0 * 2 + b) * 64 + lane];
      w3[b] = w8[(1 * 2 + b) * 64 + lane];
    }
    const dev::gptr<u32x4> mk = (dev::gptr<u32x4>)a.mk;
    u32x4 ta[4], tb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) ta[q] = mk[q * 64 + lane];

    // The four layers of one wavelet over the wave's 64 pixels.  The output
    // layer's variant b carries W4 in rows 4b..4b+2 and takes group b, so
    // lane l ends with pixel base + l's outputs in elements 0..2.
    // Software-pipelined form (PIPE): the two pixel groups go through the
    // layers one stage apart, and every stage pairs one group's MFMAs with
    // the other group's ReLU + convert:
    //   M1(0) | M1(1) R(0) | M2(0) R(1) | M2(1) R(0) | M3(0) R(1) |
    //   M3(1) R(0) | M4(0) R(1) | M4(1)
    // Per-lane arithmetic and its order are the plain step's.
    auto step_pipe = [&](const u32x4(&t)[4]) -> f32x16 {
      const f32x16 z = {};
      f32x16 d[2][2], o = z;
      i32x8 x[2];
      auto M1 = [&](int g) {
#pragma unroll
        for (int b = 0; b < 2; ++b)
          d[g][b] = mfma_bf16(t[2 * b + 1], f[g][1], mfma_bf16(t[2 * b], f[g][0], z));
      };
      auto R = [&](int g) { x[g] = relu_fp8(d[g][0], d[g][1]); };
      auto M2 = [&](int g) {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          d[g][b] = mfma_fp8(w2[b], x[g], z, sa2, sb2);
          if (AUX) d[g][b] = mfma_bf16(waux[b], auxb[g], d[g][b]);
        }
      };
      auto M3 = [&](int g) {
#pragma unroll
        for (int b = 0; b < 2; ++b) d[g][b] = mfma_fp8(w3[b], x[g], z, sa3, sb3);
      };
      auto M4 = [&](int g) { o = mfma_fp8(w4s[64 * g + lane], x[g], o, sa4, sb4); };
      constexpr int kR = 32;  // VALU of one group's ReLU + convert
      M1(0);
      __builtin_amdgcn_sched_barrier(0);
      M1(1); R(0); dev::interleave_stage<4, kR>();
      M2(0); R(1); dev::interleave_stage<AUX ? 4 : 2, kR>();
      M2(1); R(0); dev::interleave_stage<AUX ? 4 : 2, kR>();
      M3(0); R(1); dev::interleave_stage<2, kR>();
      M3(1); R(0); dev::interleave_stage<2, kR>();
      M4(0); R(1); dev::interleave_stage<1, kR>();
      M4(1);
      return o;
    };
    auto step = [&](const u32x4(&t)[4]) -> f32x16 {
      if constexpr (MODE == 1) return step_pipe(t);
      const f32x16 z = {};
      f32x16 d[2][2];
      i32x8 x[2];
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          d[g][b] = mfma_bf16(t[2 * b + 1], f[g][1], mfma_bf16(t[2 * b], f[g][0], z));
#pragma unroll
      for (int g = 0; g < 2; ++g) x[g] = relu_fp8(d[g][0], d[g][1]);
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          d[g][b] = mfma_fp8(w2[b], x[g], z, sa2, sb2);
          if (AUX) d[g][b] = mfma_bf16(waux[b], auxb[g], d[g][b]);
        }
#pragma unroll
      for (int g = 0; g < 2; ++g) x[g] = relu_fp8(d[g][0], d[g][1]);
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int b = 0; b < 2; ++b) d[g][b] = mfma_fp8(w3[b], x[g], z, sa3, sb3);
#pragma unroll
      for (int g = 0; g < 2; ++g) x[g] = relu_fp8(d[g][0], d[g][1]);
      const f32x16 o = mfma_fp8(w4s[lane], x[0], z, sa4, sb4);
      return mfma_fp8(w4s[64 + lane], x[1], o, sa4, sb4);
    };
    // Wavelets two at a time: the clamped terms of k and k + 1 are added and
    // the pair sum enters the Kahan-compensated running sum.
    using LT = dev::LightTerm<SOFTPLUS>;
    if constexpr (MODE == 2) {
      float first[3] = {0.f, 0.f, 0.f};
      for (int k = 0; k < a.k; ++k) {
        const f32x16 z = {};
        i32x8 x[2];
#pragma unroll
        for (int g = 0; g < 2; ++g) {
          const f32x16 d0 = mfma_bf16(ta[1], f[g][1], mfma_bf16(ta[0], f[g][0], z));
          const f32x16 d1 = mfma_bf16(ta[3], f[g][1], mfma_bf16(ta[2], f[g][0], z));
          x[g] = relu_fp8(d0, d1);
        }
        if (k + 1 < a.k) {
#pragma unroll
          for (int q = 0; q < 4; ++q) ta[q] = mk[((size_t)(k + 1) * 4 + q) * 64 + lane];
        }
        f32x16 o = z;
#pragma unroll
        for (int g = 0; g < 2; ++g) {
          f32x16 d0 = mfma_fp8(w2[0], x[g], z, sa2, sb2);
          if (AUX) d0 = mfma_bf16(waux[0], auxb[g], d0);
          f32x16 d1 = mfma_fp8(w2[1], x[g], z, sa2, sb2);
          if (AUX) d1 = mfma_bf16(waux[1], auxb[g], d1);
          i32x8 y = relu_fp8(d0, d1);
          d0 = mfma_fp8(w3[0], y, z, sa3, sb3);
          d1 = mfma_fp8(w3[1], y, z, sa3, sb3);
          y = relu_fp8(d0, d1);
          o = mfma_fp8(w4s[64 * g + lane], y, o, sa4, sb4);
        }
        const bool odd = (k & 1) != 0;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          const float l = LT::light(dev::load_invariant(a.light + 3 * k + ch));
          if (odd) {
            dev::kahan_add(sum[ch], comp[ch], pair_sum<SOFTPLUS>(first[ch], l, o[ch]));
          } else {
            first[ch] = LT::term(l, o[ch]);
            if (k + 1 == a.k) dev::kahan_add(sum[ch], comp[ch], first[ch]);
          }
          if (DUMP) {
            const int q = my_px - a.dump_begin;
            if (mine && q >= 0 && q < a.dump_count)
              a.dump[((size_t)q * a.k + k) * 3 + ch] = dev::transport<SOFTPLUS>(o[ch]);
          }
        }
      }
    }
    for (int k = 0; MODE != 2 && k < a.k; k += 2) {
      const bool two = k + 1 < a.k;
      if (two) {
#pragma unroll
        for (int q = 0; q < 4; ++q) tb[q] = mk[((size_t)(k + 1) * 4 + q) * 64 + lane];
      }
      const f32x16 o0 = step(ta);
      // Keep the two wavelets' layer chains apart: interleaved, their
      // accumulators (64 VGPRs each) would not fit beside the operands.
      __builtin_amdgcn_sched_barrier(0);
      f32x16 o1 = {};
      if (two) {
        if (k + 2 < a.k) {
#pragma unroll
          for (int q = 0; q < 4; ++q) ta[q] = mk[((size_t)(k + 2) * 4 + q) * 64 + lane];
        }
        o1 = step(tb);
      }
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        float x = LT::term(LT::light(dev::load_invariant(a.light + 3 * k + ch)), o0[ch]);
        if (two) x = pair_sum<SOFTPLUS>(x, LT::light(dev::load_invariant(a.light + 3 * (k + 1) + ch)), o1[ch]);
        dev::kahan_add(sum[ch], comp[ch], x);
        if (DUMP) {
          const int q = my_px - a.dump_begin;
          if (mine && q >= 0 && q < a.dump_count) {
            a.dump[((size_t)q * a.k + k) * 3 + ch] = dev::transport<SOFTPLUS>(o0[ch]);
            if (two)
              a.dump[((size_t)q * a.k + k + 1) * 3 + ch] = dev::transport<SOFTPLUS>(o1[ch]);
          }
        }
      }
    }
  }
  if (mine) {
    constexpr float f = dev::LightTerm<SOFTPLUS>::kSumScale;
    dev::write_outputs(a, my_px, hit ? (sum[0] - comp[0]) * f : 0.f,
                       hit ? (sum[1] - comp[1]) * f : 0.f, hit ? (sum[2] - comp[2]) * f : 0.f);
  }
}

// The wave's copy of the output layer's two tiles (lane l: entries l, 64 + l).
__device__ __forceinline__ void load_w4(const RelightArgs &a, i32x8 *w4s, int lane) {
  const dev::gptr<i32x8> src = (dev::gptr<i32x8>)a.w8_tiles + 2 * 2 * 64;
  w4s[lane] = src[lane];
  w4s[64 + lane] = src[64 + lane];
}

// Grid: ceil(P / 256) workgroups of 4 waves; wave w of the workgroup with
// tile t = xcd_tile(b) owns pixels [64 (4 t + w), +64) (one contiguous band
// per XCD).
template <bool SOFTPLUS, bool DUMP, bool AUX, int MODE>
__global__ __launch_bounds__(kBlock, MODE == 2 ? 3 : 2) void relight_fp8_kernel(RelightArgs a) {
  __shared__ i32x8 w4s[4][2 * 64];
  __shared__ uint32_t feat_lds[AUX ? 1 : 4 * 64 * 16];  // fused encode: [wave][pixel][16]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int base = (dev::xcd_tile(blockIdx.x, gridDim.x) * 4 + wave) * 64;
  if (base >= a.n_pixels) return;
  load_w4(a, w4s[wave], lane);
  relight_fp8_wave<SOFTPLUS, DUMP, AUX, !AUX, MODE>(a, base, lane, w4s[wave],
                                                    feat_lds + (AUX ? 0 : wave * 64 * 16));
}

// Persistent batch form (as relight_batch_kernel): a machine-sized grid
// whose waves take (job, 64-pixel group) items from a shared counter.  Each
// item reloads its job's output-layer tiles into the wave's own LDS copy.
// HAS_AUX: jobs with aux inputs take the aux path (their features come from
// the standalone encode); the others use the fused encode.
template <bool SOFTPLUS, bool HAS_AUX, int MODE>
__global__ __launch_bounds__(kBlock, MODE == 2 ? 3 : 2) void relight_fp8_batch_kernel(
    const RelightArgs *__restrict__ jobs, const BatchItem *__restrict__ items,
    const int *__restrict__ n_items_p, unsigned int *__restrict__ counter) {
  __shared__ i32x8 w4s[4][2 * 64];
  __shared__ uint32_t feat_lds[4 * 64 * 16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t *my_lds = feat_lds + wave * 64 * 16;
  const int n_items = *n_items_p;
  auto fetch = [&]() {
    int it = 0;
    if (lane == 0) it = (int)atomicAdd(counter, 1u);
    return __builtin_amdgcn_readfirstlane(__shfl(it, 0));
  };
  int next = fetch();
  for (;;) {
    const int it = next;
    if (it >= n_items) break;
    next = fetch();
    const BatchItem item = items[it];
    const RelightArgs a = jobs[item.job];
    load_w4(a, w4s[wave], lane);
    if (HAS_AUX && a.aux)
      relight_fp8_wave<SOFTPLUS, false, true, false, MODE>(a, item.base, lane, w4s[wave], my_lds);
    else
      relight_fp8_wave<SOFTPLUS, false, false, true, MODE>(a, item.base, lane, w4s[wave], my_lds);
  }
}

struct Fp8Variant {
  const char *name;
};
constexpr Fp8Variant kFp8Variants[] = {{"fp8.l1bf16.occ2"}, {"fp8.pipe.occ2"}, {"fp8.seq.occ3"}};
constexpr int kNumFp8Variants = sizeof(kFp8Variants) / sizeof(kFp8Variants[0]);

// Exponent e with amax * 2^e in (224, 448], 0 for a zero or non-finite amax.
int weight_exp(float amax) {
  if (!(amax > 0.f) || !std::isfinite(amax)) return 0;
  int E = 0;
  const float m = std::frexp(amax, &E);  // amax = m 2^E, m in [0.5, 1)
  const int e = m <= 0.875f ? 9 - E : 8 - E;
  return std::min(std::max(e, -100), 100);
}

}  // namespace

uint8_t f32_to_e4m3(float x) {
  if (std::isnan(x)) return 0x7f;
  const uint8_t sign = std::signbit(x) ? 0x80 : 0;
  const float ax = std::fabs(x);
  if (ax >= 448.f) return sign | 0x7e;  // saturate
  if (ax < 0.015625f) {                 // subnormal range: multiples of 2^-9
    const float q = std::nearbyint(ax * 512.f);  // 0..8; 8 is the smallest normal
    return sign | (uint8_t)q;
  }
  int E = 0;
  std::frexp(ax, &E);  // ax in [2^(E-1), 2^E)
  const int ub = E - 1;
  int mant = (int)std::nearbyint(std::ldexp(ax, 3 - ub)) - 8;  // 0..8
  int exp = ub + 7;
  if (mant == 8) {
    mant = 0;
    ++exp;
  }
  if (exp > 15 || (exp == 15 && mant == 7)) return sign | 0x7e;
  return sign | (uint8_t)(exp << 3) | (uint8_t)mant;
}

float e4m3_to_f32(uint8_t c) {
  const int e = (c >> 3) & 15, m = c & 7;
  if (e == 15 && m == 7) return NAN;
  const float v = e == 0 ? std::ldexp((float)m, -9) : std::ldexp(1.0f + m / 8.0f, e - 7);
  return (c & 0x80) ? -v : v;
}

// Tiles [layer][b][lane][j], 32 bytes per lane: layers 2, 3 (b = 32-row
// output block) and the output layer's two variants (b: W4 in rows
// 4b..4b+2, zeros elsewhere); slot (lane / 32, j) carries input pi(h, j).
void build_fp8_weight_tiles(const float *w, std::vector<uint8_t> &out, int exps[3]) {
  constexpr int kW1 = 64 * 32, kW2 = 64 * 64, kW3 = 64 * 64;
  const float *mats[3] = {w + kW1, w + kW1 + kW2, w + kW1 + kW2 + kW3};
  const int rows[3] = {64, 64, 3};
  for (int l = 0; l < 3; ++l) {
    float amax = 0.f;
    for (int i = 0; i < rows[l] * 64; ++i) amax = std::max(amax, std::fabs(mats[l][i]));
    exps[l] = weight_exp(amax);
  }
  auto pi = [](int h, int j) { return 32 * (j >> 4) + 8 * ((j & 15) >> 2) + 4 * h + (j & 3); };
  out.assign(3 * 2 * 64 * 32, 0);
  for (int l = 0; l < 3; ++l) {
    const float s = std::ldexp(1.0f, exps[l]);
    for (int b = 0; b < 2; ++b)
      for (int lane = 0; lane < 64; ++lane)
        for (int j = 0; j < 32; ++j) {
          const int r = lane & 31, in = pi(lane >> 5, j);
          float v = 0.f;
          if (l < 2) v = mats[l][(32 * b + r) * 64 + in];
          else if (r >= 4 * b && r < 4 * b + 3) v = mats[2][(r - 4 * b) * 64 + in];
          out[(((size_t)l * 2 + b) * 64 + lane) * 32 + j] = f32_to_e4m3(v * s);
        }
  }
}

int relight_fp8_variant_count() { return kNumFp8Variants; }
const char *relight_fp8_variant_name(int v) {
  return v >= 0 && v < kNumFp8Variants ? kFp8Variants[v].name : nullptr;
}

void launch_relight_fp8(const RelightArgs &a, int variant, hipStream_t s) {
  require(a.n_pixels >= 1 && a.k >= 1 && a.k <= NWPRT_MAX_K, NWPRT_ERR_INVALID_ARGUMENT,
          "relight: bad pixel count or K");
  require((a.fused_encode ? a.position && a.hash : a.features != nullptr) && a.mk && a.light &&
              a.w8_tiles && a.amax && a.direct,
          NWPRT_ERR_INVALID_ARGUMENT, "relight: missing FP8 operand");
  require(!a.aux || (a.waux32_tiles && !a.fused_encode), NWPRT_ERR_INVALID_ARGUMENT,
          "relight: missing FP8 aux weights");
  require(variant >= 0 && variant < kNumFp8Variants, NWPRT_ERR_INVALID_ARGUMENT,
          "relight: unknown FP8 kernel variant");
  const dim3 grid((a.n_pixels + kBlock - 1) / kBlock), block(kBlock);
  const bool aux = a.aux != nullptr, sp = a.softplus != 0, dump = a.dump != nullptr;
  auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, grid, block, 0, s, a); };
#define NWPRT_FP8_PICK(SP, DUMP, AUX)                            \
  if (variant == 1) go(relight_fp8_kernel<SP, DUMP, AUX, 1>);    \
  else if (variant == 2) go(relight_fp8_kernel<SP, DUMP, AUX, 2>); \
  else go(relight_fp8_kernel<SP, DUMP, AUX, 0>);
  if (sp) {
    if (aux) {
      if (dump) { NWPRT_FP8_PICK(true, true, true) } else { NWPRT_FP8_PICK(true, false, true) }
    } else {
      if (dump) { NWPRT_FP8_PICK(true, true, false) } else { NWPRT_FP8_PICK(true, false, false) }
    }
  } else {
    if (aux) {
      if (dump) { NWPRT_FP8_PICK(false, true, true) } else { NWPRT_FP8_PICK(false, false, true) }
    } else {
      if (dump) { NWPRT_FP8_PICK(false, true, false) } else { NWPRT_FP8_PICK(false, false, false) }
    }
  }
#undef NWPRT_FP8_PICK
  hip_check(hipGetLastError(), "relight_fp8_kernel launch");
}

void launch_relight_fp8_batch(const RelightArgs *jobs, const BatchItem *items, const int *n_items,
                              unsigned int *counter, bool softplus, bool has_aux, int variant,
                              int n_cu, hipStream_t s) {
  require(jobs && items && n_items && counter && n_cu >= 1, NWPRT_ERR_INVALID_ARGUMENT,
          "batch: missing operand");
  require(variant >= 0 && variant < kNumFp8Variants, NWPRT_ERR_INVALID_ARGUMENT,
          "batch: unknown FP8 kernel variant");
  // A machine-sized grid: as many 4-wave workgroups per CU as the form
  // keeps resident (two, three for the seq form).
  const dim3 grid(n_cu * (variant == 2 ? 3 : 2)), block(kBlock);
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, grid, block, 0, s, jobs, items, n_items, counter);
  };
#define NWPRT_FP8B_PICK(SP, AUX)                               \
  if (variant == 1) go(relight_fp8_batch_kernel<SP, AUX, 1>);  \
  else if (variant == 2) go(relight_fp8_batch_kernel<SP, AUX, 2>); \
  else go(relight_fp8_batch_kernel<SP, AUX, 0>);
  if (softplus) {
    if (has_aux) { NWPRT_FP8B_PICK(true, true) } else { NWPRT_FP8B_PICK(true, false) }
  } else {
    if (has_aux) { NWPRT_FP8B_PICK(false, true) } else { NWPRT_FP8B_PICK(false, false) }
  }
#undef NWPRT_FP8B_PICK
  hip_check(hipGetLastError(), "relight_fp8_batch_kernel launch");
}

// A tiles of the second layer's 16 aux columns for the FP8 form's BF16
// v_mfma_f32_32x32x16: tile [b][lane][s] = Waux[32 b + lane%32][8 (lane/32) + s].
void build_fp8_aux_tiles(const float *waux, std::vector<uint16_t> &out) {
  out.assign(2 * 64 * 8, 0);
  for (int b = 0; b < 2; ++b)
    for (int lane = 0; lane < 64; ++lane)
      for (int s = 0; s < 8; ++s)
        out[((size_t)b * 64 + lane) * 8 + s] =
            f32_to_bf16_bits(waux[(32 * b + (lane & 31)) * 16 + 8 * (lane >> 5) + s]);
}

}  // namespace nwprt
