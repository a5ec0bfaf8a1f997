// Per-(pixel, wavelet) network and per-pixel reduction for gfx950.
//
// Reference path being replaced (Renderer/tinycudann/tcnn.cu:111-152): a host
// loop over image rows runs, per row, the hash grid, the per-pair Hadamard
// product with the wavelet feature (utils_kernels.h:97-115), the out_prd and
// cmp_net networks and a cast, materialising 16 + 3 floats per pair; then
// computeResponse (utils_kernels.h:51-90) sums max(light_i * T_i, 0) over K.
//
// Here one launch does all pairs of a frame and nothing per pair touches
// HBM: each wave owns 64 pixels and walks the K selected wavelets; the four
// linear layers run as chained 16x16x32 MFMAs whose results stay in VGPRs,
// and the clamped sum is carried per lane with Kahan compensation.
//
// The wave-level code (MFMA layout, layers, fused encode, per-wave relight)
// is in relight_mma16_device.hpp.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.hpp"
#include "hashgrid_device.hpp"
#include "kernels.hpp"
#include "relight_device.hpp"
#include "relight_mma16_device.hpp"

namespace nwprt {
namespace {

using dev::f32x4;
using dev::u32x4;

constexpr int kRelightBlock = 256;  // 4 waves x 64 pixels

using namespace mma16;

// Tile-shape variants of the fused kernel (chosen per shape by the
// autotuner): GP pixel groups carried through the layers together (GP
// independent MFMA chains hide each other's result latency), the output
// layer's A tiles in VGPRs or in LDS (LDS frees 32 VGPRs for a third wave per
// SIMD), and the minimum waves per SIMD the register budget must allow.
struct MfmaVariant {
  int gp;
  bool w4_lds;
  int min_waves;
  const char *name;
};
constexpr MfmaVariant kMfmaVariants[] = {
    {2, false, 2, "gp2.w4reg.occ2"}, {4, false, 2, "gp4.w4reg.occ2"},
    {2, true, 3, "gp2.w4lds.occ3"},  {4, true, 2, "gp4.w4lds.occ2"},
    {1, true, 3, "gp1.w4lds.occ3"},  {0, false, 2, "pipe.w4reg.occ2"},
    {0, true, 2, "pipe.w4lds.occ2"},
};
constexpr int kNumMfmaVariants = sizeof(kMfmaVariants) / sizeof(kMfmaVariants[0]);

// Grid: ceil(P / 256) workgroups of 4 waves; wave w of the workgroup with
// tile t = xcd_tile(b) owns pixels [64 (4 t + w), +64), so each XCD relights
// (and with the fused encode, gathers hash entries for) one contiguous band.
template <typename E, bool SOFTPLUS, bool DUMP, bool AUX, int GP, bool W4LDS, int MINW>
__global__ __launch_bounds__(kRelightBlock, MINW) void relight_mfma_kernel(RelightArgs a) {
  __shared__ u32x4 w4s[W4LDS ? 4 * 2 * 64 : 1];
  __shared__ uint32_t feat_lds[AUX ? 1 : 4 * 64 * 16];  // fused encode: [wave][pixel][16]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (W4LDS) {
    const u32x4 *wt = (const u32x4 *)a.w_tiles;
    for (int i = threadIdx.x; i < 4 * 2 * 64; i += kRelightBlock) w4s[i] = wt[2 * 8 * 64 + i];
    __syncthreads();
  }
  const int base = (dev::xcd_tile(blockIdx.x, gridDim.x) * 4 + wave) * 64;
  if (base >= a.n_pixels) return;
  relight_wave<E, SOFTPLUS, DUMP, AUX, GP, W4LDS, !AUX>(a, base, lane, w4s,
                                                       feat_lds + (AUX ? 0 : wave * 64 * 16));
}

// Persistent batch form: a machine-sized grid whose waves take work items
// (job, first pixel of a 64-pixel group) from a shared counter until the
// list is exhausted.  Jobs are whole relights with their own frame,
// scene weights, selection and K; items are listed by decreasing K so the
// tail of the launch is short items.  Output-layer tiles stay in VGPRs, so
// there is no workgroup barrier and no scene switch cost beyond reloading;
// jobs with the fused encode use their wave's own 4 KB of LDS.  HAS_AUX
// compiles in the path of jobs with aux inputs (features and aux inputs from
// the standalone encode).
template <typename E, bool SOFTPLUS, bool HAS_AUX, int GP, int MINW>
__global__ __launch_bounds__(kRelightBlock, MINW) void relight_batch_kernel(
    const RelightArgs *__restrict__ jobs, const BatchItem *__restrict__ items,
    const int *__restrict__ n_items_p, unsigned int *__restrict__ counter) {
  __shared__ uint32_t feat_lds[4 * 64 * 16];  // fused encode: [wave][pixel][16]
  const int lane = threadIdx.x & 63;
  uint32_t *my_lds = feat_lds + (threadIdx.x >> 6) * 64 * 16;
  const int n_items = *n_items_p;
  // The next item's index is fetched while the current one runs, so the
  // counter's round trip is off the critical path.
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
    // One 64-pixel group per item: a loop over several would let the
    // compiler hoist the group's loads and spill.  Jobs with aux inputs take
    // the tile shape their single relights use (two groups at a time).
    if (HAS_AUX && a.aux)
      relight_wave<E, SOFTPLUS, false, true, 2, false, false>(a, item.base, lane, nullptr, my_lds);
    else
      relight_wave<E, SOFTPLUS, false, false, GP, false, true>(a, item.base, lane, nullptr,
                                                               my_lds);
  }
}

__global__ void reset_counter_kernel(unsigned int *counter) { *counter = 0u; }

// Plain FP32 form: one lane per pixel walks the K wavelets with the four
// layers on the vector ALU, weights broadcast from LDS and the selected
// wavelet features read with uniform (scalar) loads.
constexpr int kPlainBlock = 256;
constexpr int kW1 = 64 * 32, kW2 = 64 * 64, kW3 = 64 * 64, kW4 = 3 * 64;

constexpr int kWaux = 64 * 16;

template <bool DUMP, bool AUX>
__global__ __launch_bounds__(kPlainBlock) void relight_plain_kernel(RelightArgs a) {
  __shared__ float w[kW1 + kW2 + kW3 + kW4 + (AUX ? kWaux : 0)];
  for (int i = threadIdx.x; i < kW1 + kW2 + kW3 + kW4; i += kPlainBlock) w[i] = a.w_f32[i];
  if (AUX)
    for (int i = threadIdx.x; i < kWaux; i += kPlainBlock) w[kW1 + kW2 + kW3 + kW4 + i] = a.waux_f32[i];
  __syncthreads();
  const int p = blockIdx.x * kPlainBlock + threadIdx.x;
  if (p >= a.n_pixels) return;
  const bool hit = a.env_flag == nullptr || a.env_flag[p] != 0;
  float sum[3] = {0.f, 0.f, 0.f}, comp[3] = {0.f, 0.f, 0.f}, pend[3] = {0.f, 0.f, 0.f};
  if (hit) {
    float h[32];
    const float4 *fp = (const float4 *)((const float *)a.features + (size_t)p * 32);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float4 v = fp[q];
      h[4 * q] = v.x; h[4 * q + 1] = v.y; h[4 * q + 2] = v.z; h[4 * q + 3] = v.w;
    }
    // Aux inputs enter the second layer as a per-pixel constant term.
    float b2[64];
#pragma unroll
    for (int o = 0; o < 64; ++o) b2[o] = 0.f;
    if (AUX) {
      const float *ax = (const float *)a.aux + (size_t)p * 16;
      const float *Wa = w + kW1 + kW2 + kW3 + kW4;
#pragma unroll
      for (int o = 0; o < 64; ++o) {
        float acc = 0.f;
#pragma unroll
        for (int i = 0; i < 16; ++i) acc = fmaf(Wa[o * 16 + i], ax[i], acc);
        b2[o] = acc;
      }
    }
    for (int k = 0; k < a.k; ++k) {
      // The weight reads are loop-invariant; an opaque zero offset keeps the
      // compiler from hoisting all ~10K of them out of the K loop.
      int wo = 0;
      asm volatile("" : "+s"(wo));
      const float *W1 = w + wo, *W2 = W1 + kW1, *W3 = W2 + kW2, *W4 = W3 + kW3;
      const float *ws = a.wsel + (size_t)k * 32;
      float y1[64];
#pragma unroll
      for (int o = 0; o < 64; ++o) {
        float acc = 0.f;
#pragma unroll
        for (int i = 0; i < 32; ++i) acc = fmaf(W1[o * 32 + i], h[i] * ws[i], acc);
        y1[o] = fmaxf(acc, 0.f);
      }
      float y2[64];
#pragma unroll
      for (int o = 0; o < 64; ++o) {
        float acc = b2[o];
#pragma unroll
        for (int i = 0; i < 64; ++i) acc = fmaf(W2[o * 64 + i], y1[i], acc);
        y2[o] = fmaxf(acc, 0.f);
      }
#pragma unroll
      for (int o = 0; o < 64; ++o) {
        float acc = 0.f;
#pragma unroll
        for (int i = 0; i < 64; ++i) acc = fmaf(W3[o * 64 + i], y2[i], acc);
        y1[o] = fmaxf(acc, 0.f);  // reuse y1 as layer-3 output
      }
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        float acc = 0.f;
#pragma unroll
        for (int i = 0; i < 64; ++i) acc = fmaf(W4[ch * 64 + i], y1[i], acc);
        const float t = a.softplus ? dev::softplus(acc) : acc;
        // Pairs of wavelets: the clamped terms of k and k + 1 are added and
        // the pair sum enters the compensated running sum (the fused kernel's
        // order).
        const float term = dev::relu_f32(a.light[3 * k + ch] * t);
        if (k & 1) dev::neumaier_add(sum[ch], comp[ch], pend[ch] + term);
        else pend[ch] = term;
        if (DUMP) {
          const int q = p - a.dump_begin;
          if (q >= 0 && q < a.dump_count) a.dump[((size_t)q * a.k + k) * 3 + ch] = t;
        }
      }
    }
    if (a.k & 1)
      for (int ch = 0; ch < 3; ++ch) dev::neumaier_add(sum[ch], comp[ch], pend[ch]);
  }
  write_outputs(a, p, hit ? sum[0] + comp[0] : 0.f, hit ? sum[1] + comp[1] : 0.f,
                hit ? sum[2] + comp[2] : 0.f);
}

// Activation calibration for the FP8 form: the BF16 network over a pixel
// subsample (a.features, a.n_pixels) and the wavelets [k0, k0 + k_chunk) of
// blockIdx.y, recording the maximum pre-activation output of layers 1..3
// (the ReLU outputs share it when positive).  One atomicMax per wave and
// layer on the float bits, which order like the values for non-negative
// floats; the maxima start at 0.
template <bool AUX>
__global__ __launch_bounds__(kRelightBlock) void calib_amax_kernel(RelightArgs a, int k_chunk) {
  using E = __bf16;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, n = lane & 15;
  const int base = (blockIdx.x * 4 + wave) * 64;
  const int k0 = blockIdx.y * k_chunk, k1 = min(a.k, k0 + k_chunk);
  const int P = a.n_pixels;
  if (base >= P || k0 >= k1) return;
  u32x4 hb[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int p = base + 16 * j + n;
    hb[j] = p < P ? dev::load_features8(a, P, p, 8 * g) : u32x4{0u, 0u, 0u, 0u};
  }
  u32x2 auxb[4], waux[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int p = base + 16 * j + n;
    auxb[j] = AUX && p < P ? *(const u32x2 *)((const uint16_t *)a.aux + (size_t)p * 16 + 4 * g)
                           : u32x2{0u, 0u};
    waux[j] = AUX ? ((const u32x2 *)a.waux_tiles)[j * 64 + lane] : u32x2{0u, 0u};
  }
  const u32x4 *wt = (const u32x4 *)a.w_tiles;
  u32x4 w2[4][2], w3[4][2];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      w2[b][kb] = wt[(0 * 8 + b * 2 + kb) * 64 + lane];
      w3[b][kb] = wt[(1 * 8 + b * 2 + kb) * 64 + lane];
    }
  const u32x4 *mk = (const u32x4 *)a.mk;
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  float m[3] = {0.f, 0.f, 0.f};
  auto upd = [](float &mm, f32x4 c) { mm = fmaxf(mm, fmaxf(fmaxf(c[0], c[1]), fmaxf(c[2], c[3]))); };
  for (int k = k0; k < k1; ++k) {
    u32x4 a1[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) a1[b] = mk[((size_t)k * 4 + b) * 64 + lane];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      f32x4 c[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        c[b] = Mfma<E>::run(a1[b], hb[j], zero);
        upd(m[0], c[b]);
      }
      u32x4 x0 = relu_pack<E>(c[0], c[1]), x1 = relu_pack<E>(c[2], c[3]);
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        c[b] = Mfma<E>::run(w2[b][1], x1, Mfma<E>::run(w2[b][0], x0, zero));
        if (AUX) c[b] = Mfma16<E>::run(waux[b], auxb[j], c[b]);
        upd(m[1], c[b]);
      }
      x0 = relu_pack<E>(c[0], c[1]);
      x1 = relu_pack<E>(c[2], c[3]);
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        c[b] = Mfma<E>::run(w3[b][1], x1, Mfma<E>::run(w3[b][0], x0, zero));
        upd(m[2], c[b]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float v = m[i];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    if (lane == 0) atomicMax((unsigned int *)a.amax + i, __float_as_uint(v));
  }
}

// Calibration subsample: calibration pixel q is frame pixel q * stride, its
// features copied in the frame's layout (zeros for a background pixel, as
// the encode pass writes them) and, with aux inputs, its 16 aux values
// (zeros for a background pixel, whose activations the relight discards);
// block 0 also clears the four maxima.  Row layout: thread i copies 16 B
// part i % 4 of pixel i / 4 (and 8 B part i % 4 of its aux row); planar:
// 8 B of plane i % 8 of pixel i / 8.
__global__ __launch_bounds__(256) void calib_gather_kernel(const uint16_t *__restrict__ feat,
                                                           const uint16_t *__restrict__ aux,
                                                           const uint8_t *__restrict__ flag,
                                                           int n_pixels, int stride,
                                                           uint16_t *__restrict__ calib,
                                                           uint16_t *__restrict__ calib_aux,
                                                           int n_cal, float *__restrict__ amax,
                                                           int planar) {
  if (blockIdx.x == 0 && threadIdx.x < 4) amax[threadIdx.x] = 0.f;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int shift = planar ? 3 : 2;
  const int q = i >> shift, part = i & ((1 << shift) - 1);
  if (q >= n_cal) return;
  const long long p = (long long)q * stride;
  const bool hit = p < n_pixels && (flag == nullptr || flag[p] != 0);
  if (planar) {
    const uint2 *src = (const uint2 *)feat;
    ((uint2 *)calib)[(size_t)part * n_cal + q] =
        hit ? src[(size_t)part * n_pixels + p] : make_uint2(0u, 0u);
  } else {
    const u32x4 v = hit ? *(const u32x4 *)(feat + p * 32 + 8 * part) : u32x4{0u, 0u, 0u, 0u};
    *(u32x4 *)(calib + (size_t)q * 32 + 8 * part) = v;
    if (aux)
      *(uint2 *)(calib_aux + (size_t)q * 16 + 4 * part) =
          hit ? *(const uint2 *)(aux + p * 16 + 4 * part) : make_uint2(0u, 0u);
  }
}

// The calibration subsample encoded from positions: thread q computes the 16
// levels of frame pixel q * stride with the standalone planar encode's
// arithmetic (plane r = levels r and 15 - r, each accumulated from zero) and
// writes the planar layout [8][n_cal][4] in BF16; background pixels get
// zeros.  Block 0 clears the four maxima.
template <bool HF16>
__global__ __launch_bounds__(256) void calib_encode_kernel(RelightArgs a, int stride,
                                                           uint16_t *__restrict__ calib,
                                                           int n_cal, float *__restrict__ amax) {
  if (blockIdx.x == 0 && threadIdx.x < 4) amax[threadIdx.x] = 0.f;
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= n_cal) return;
  const long long p = (long long)q * stride;
  const bool hit = p < a.n_pixels && (a.env_flag == nullptr || a.env_flag[p] != 0);
  float x[3] = {0.f, 0.f, 0.f};
  if (hit) {
    const float4 x4 = *(const float4 *)(a.position + 4 * (size_t)p);
    x[0] = x4.x;
    x[1] = x4.y;
    x[2] = x4.z;
  }
  uint2 *dst = (uint2 *)calib;
  for (int r = 0; r < 8; ++r) {
    float v[4] = {0.f, 0.f, 0.f, 0.f};
    if (hit) {
      dev::accumulate_level<2, HF16>(a.lv, a.hash, r, x, v, nullptr);
      dev::accumulate_level<2, HF16>(a.lv, a.hash, 15 - r, x, v + 2, nullptr);
    }
    dst[(size_t)r * n_cal + q] = make_uint2(dev::pack_bf16(v[0], v[1]), dev::pack_bf16(v[2], v[3]));
  }
}

__global__ __launch_bounds__(256) void tonemap_kernel(const float *__restrict__ src,
                                                      uint32_t *__restrict__ dst, int n) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const float4 v = *(const float4 *)(src + 4 * (size_t)p);
  dst[p] = dev::tonemap_rgba8(v.x, v.y, v.z);
}

template <typename E, bool SP, int V>
void launch_v(const RelightArgs &a, dim3 grid, dim3 block, hipStream_t s) {
  constexpr MfmaVariant v = kMfmaVariants[V];
  if (a.aux) {
    // Aux scenes are built for variant 0's tile shape (tap included).
    if (a.dump)
      hipLaunchKernelGGL((relight_mfma_kernel<E, SP, true, true, 2, false, 2>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((relight_mfma_kernel<E, SP, false, true, 2, false, 2>), grid, block, 0, s,
                         a);
  } else if (a.dump && V == 0) {
    hipLaunchKernelGGL((relight_mfma_kernel<E, SP, true, false, v.gp, v.w4_lds, v.min_waves>),
                       grid, block, 0, s, a);
  } else {
    hipLaunchKernelGGL((relight_mfma_kernel<E, SP, false, false, v.gp, v.w4_lds, v.min_waves>),
                       grid, block, 0, s, a);
  }
}

template <typename E, bool SP>
void launch_sp(const RelightArgs &a, int variant, hipStream_t s) {
  const dim3 grid((a.n_pixels + kRelightBlock - 1) / kRelightBlock), block(kRelightBlock);
  if (a.dump) variant = 0;  // the transport tap is built for the default variant only
  switch (variant) {
    case 0: launch_v<E, SP, 0>(a, grid, block, s); break;
    case 1: launch_v<E, SP, 1>(a, grid, block, s); break;
    case 2: launch_v<E, SP, 2>(a, grid, block, s); break;
    case 3: launch_v<E, SP, 3>(a, grid, block, s); break;
    case 4: launch_v<E, SP, 4>(a, grid, block, s); break;
    case 5: launch_v<E, SP, 5>(a, grid, block, s); break;
    case 6: launch_v<E, SP, 6>(a, grid, block, s); break;
    default: fail(NWPRT_ERR_INVALID_ARGUMENT, "relight: unknown kernel variant");
  }
}

template <typename E, bool SP, int V>
void launch_batch_v(const RelightArgs *jobs, const BatchItem *items, const int *n_items,
                    unsigned int *counter, bool has_aux, int n_cu, hipStream_t s) {
  constexpr MfmaVariant v = kMfmaVariants[V];
  static_assert(!v.w4_lds, "the batch form keeps the output layer in VGPRs");
  if (has_aux)
    hipLaunchKernelGGL((relight_batch_kernel<E, SP, true, v.gp, v.min_waves>),
                       dim3(n_cu * v.min_waves), dim3(kRelightBlock), 0, s, jobs, items, n_items,
                       counter);
  else
    hipLaunchKernelGGL((relight_batch_kernel<E, SP, false, v.gp, v.min_waves>),
                       dim3(n_cu * v.min_waves), dim3(kRelightBlock), 0, s, jobs, items, n_items,
                       counter);
}

template <typename E, bool SP>
void launch_batch_e(const RelightArgs *jobs, const BatchItem *items, const int *n_items,
                    unsigned int *counter, bool has_aux, int variant, int n_cu, hipStream_t s) {
  switch (variant) {
    case 0: launch_batch_v<E, SP, 0>(jobs, items, n_items, counter, has_aux, n_cu, s); break;
    case 1: launch_batch_v<E, SP, 1>(jobs, items, n_items, counter, has_aux, n_cu, s); break;
    case 5: launch_batch_v<E, SP, 5>(jobs, items, n_items, counter, has_aux, n_cu, s); break;
    default: fail(NWPRT_ERR_INVALID_ARGUMENT, "batch: variant has no persistent form");
  }
}

template <typename E>
void launch_mfma_variant(const RelightArgs &a, int variant, hipStream_t s) {
  if (a.softplus) launch_sp<E, true>(a, variant, s);
  else launch_sp<E, false>(a, variant, s);
}

}  // namespace

// A-operand tiles of layers 2, 3 (64 x 64) and the output layer (3 x 64) for
// the chained MFMA layout above.  Tile [b][kb][lane][s] holds
// W[16 b + lane%16][f(kb, lane/16, s)] with the contraction permutation
// f(kb, g, s) = 32 kb + (s < 4 ? 4 g + s : 16 + 4 g + s - 4).  The output
// layer is stored as four variants j (in place of b): rows 4j..4j+2 hold W4,
// the other rows are zero.
void build_weight_tiles(const float *w, Elem elem, std::vector<uint16_t> &out) {
  const float *W2 = w + kW1, *W3 = W2 + kW2, *W4 = W3 + kW3;
  out.assign(3 * 4 * 2 * 64 * 8, 0);
  auto cvt = [&](float v) {
    return elem == Elem::BF16 ? f32_to_bf16_bits(v) : f32_to_f16_bits(v);
  };
  auto feat = [](int kb, int g, int s) { return 32 * kb + (s < 4 ? 4 * g + s : 16 + 4 * g + s - 4); };
  for (int layer = 0; layer < 3; ++layer)
    for (int b = 0; b < 4; ++b)
      for (int kb = 0; kb < 2; ++kb)
        for (int lane = 0; lane < 64; ++lane)
          for (int s = 0; s < 8; ++s) {
            const int row = lane & 15, g = lane >> 4, f = feat(kb, g, s);
            float v = 0.f;
            if (layer == 0) v = W2[(16 * b + row) * 64 + f];
            else if (layer == 1) v = W3[(16 * b + row) * 64 + f];
            else if (row >= 4 * b && row < 4 * b + 3) v = W4[(row - 4 * b) * 64 + f];
            out[((((size_t)layer * 4 + b) * 2 + kb) * 64 + lane) * 8 + s] = cvt(v);
          }
}

int relight_mfma_variant_count() { return kNumMfmaVariants; }
const char *relight_mfma_variant_name(int v) {
  return v >= 0 && v < kNumMfmaVariants ? kMfmaVariants[v].name : nullptr;
}

// A tiles of the second layer's 16 aux columns for v_mfma_f32_16x16x16:
// tile [b][lane][s] = Waux[16 b + lane%16][4 (lane/16) + s].
void build_aux_tiles(const float *waux, Elem elem, std::vector<uint16_t> &out) {
  out.assign(4 * 64 * 4, 0);
  for (int b = 0; b < 4; ++b)
    for (int lane = 0; lane < 64; ++lane)
      for (int s = 0; s < 4; ++s) {
        const float v = waux[(16 * b + (lane & 15)) * 16 + 4 * (lane >> 4) + s];
        out[((size_t)b * 64 + lane) * 4 + s] =
            elem == Elem::BF16 ? f32_to_bf16_bits(v) : f32_to_f16_bits(v);
      }
}

void launch_relight_mfma(const RelightArgs &a, Elem elem, int variant, hipStream_t s) {
  require(a.n_pixels >= 1 && a.k >= 1 && a.k <= NWPRT_MAX_K, NWPRT_ERR_INVALID_ARGUMENT,
          "relight: bad pixel count or K");
  require((a.fused_encode ? a.position && a.hash : a.features != nullptr) && a.mk && a.light &&
              a.w_tiles && a.direct,
          NWPRT_ERR_INVALID_ARGUMENT, "relight: missing operand");
  require(!a.aux || a.waux_tiles, NWPRT_ERR_INVALID_ARGUMENT, "relight: missing aux weights");
  require(!a.fused_encode || (!a.aux && a.lv.n_levels == 16), NWPRT_ERR_INVALID_ARGUMENT,
          "relight: the fused encode needs 16 x 2 grids without aux inputs");
  if (elem == Elem::BF16) {
    launch_mfma_variant<__bf16>(a, variant, s);
  } else if (elem == Elem::F16) {
    launch_mfma_variant<_Float16>(a, variant, s);
  } else {
    fail(NWPRT_ERR_INVALID_ARGUMENT, "relight: MFMA element must be BF16 or FP16");
  }
  hip_check(hipGetLastError(), "relight_mfma_kernel launch");
}

void launch_relight_plain(const RelightArgs &a, hipStream_t s) {
  require(a.n_pixels >= 1 && a.k >= 1 && a.k <= NWPRT_MAX_K, NWPRT_ERR_INVALID_ARGUMENT,
          "relight: bad pixel count or K");
  require(a.features && a.wsel && a.light && a.w_f32 && a.direct, NWPRT_ERR_INVALID_ARGUMENT,
          "relight: missing operand");
  const dim3 grid((a.n_pixels + kPlainBlock - 1) / kPlainBlock), block(kPlainBlock);
  if (a.aux) {
    require(a.waux_f32 != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "relight: missing aux weights");
    if (a.dump) hipLaunchKernelGGL((relight_plain_kernel<true, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((relight_plain_kernel<false, true>), grid, block, 0, s, a);
  } else {
    if (a.dump) hipLaunchKernelGGL((relight_plain_kernel<true, false>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((relight_plain_kernel<false, false>), grid, block, 0, s, a);
  }
  hip_check(hipGetLastError(), "relight_plain_kernel launch");
}

bool relight_batch_variant_ok(int v) {
  return v >= 0 && v < kNumMfmaVariants && !kMfmaVariants[v].w4_lds;
}

void launch_relight_batch(const RelightArgs *jobs, const BatchItem *items, const int *n_items,
                          unsigned int *counter, Elem elem, bool softplus, bool has_aux,
                          int variant, int n_cu, hipStream_t s) {
  require(jobs && items && n_items && counter && n_cu >= 1, NWPRT_ERR_INVALID_ARGUMENT,
          "batch: missing operand");
  if (elem == Elem::BF16) {
    if (softplus)
      launch_batch_e<__bf16, true>(jobs, items, n_items, counter, has_aux, variant, n_cu, s);
    else
      launch_batch_e<__bf16, false>(jobs, items, n_items, counter, has_aux, variant, n_cu, s);
  } else if (elem == Elem::F16) {
    if (softplus)
      launch_batch_e<_Float16, true>(jobs, items, n_items, counter, has_aux, variant, n_cu, s);
    else
      launch_batch_e<_Float16, false>(jobs, items, n_items, counter, has_aux, variant, n_cu, s);
  } else {
    fail(NWPRT_ERR_INVALID_ARGUMENT, "batch: MFMA element must be BF16 or FP16");
  }
  hip_check(hipGetLastError(), "relight_batch_kernel launch");
}

void launch_reset_counter(unsigned int *counter, hipStream_t s) {
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, counter);
  hip_check(hipGetLastError(), "reset_counter_kernel launch");
}

void launch_calib_gather(const uint16_t *features, const uint16_t *aux, const uint8_t *env_flag,
                         int n_pixels, int stride, uint16_t *calib, uint16_t *calib_aux,
                         int n_cal, float *amax, bool planar, hipStream_t s) {
  require(features && calib && amax && n_pixels >= 1 && stride >= 1 && n_cal >= 1 &&
              (long long)(n_cal - 1) * stride < n_pixels && (!aux || (calib_aux && !planar)),
          NWPRT_ERR_INVALID_ARGUMENT, "calibration gather: bad subsample");
  const int threads = (planar ? 8 : 4) * n_cal;
  hipLaunchKernelGGL(calib_gather_kernel, dim3((threads + 255) / 256), dim3(256), 0, s, features,
                     aux, env_flag, n_pixels, stride, calib, calib_aux, n_cal, amax, (int)planar);
  hip_check(hipGetLastError(), "calib_gather_kernel launch");
}

void launch_calib_encode(const RelightArgs &enc, int stride, uint16_t *calib, int n_cal,
                         float *amax, hipStream_t s) {
  require(enc.position && enc.hash && calib && amax && enc.n_pixels >= 1 && stride >= 1 &&
              n_cal >= 1 && (long long)(n_cal - 1) * stride < enc.n_pixels &&
              enc.lv.n_levels == 16,
          NWPRT_ERR_INVALID_ARGUMENT, "calibration encode: bad subsample");
  const dim3 grid((n_cal + 255) / 256);
  if (enc.hash_f16)
    hipLaunchKernelGGL(calib_encode_kernel<true>, grid, dim3(256), 0, s, enc, stride, calib, n_cal,
                       amax);
  else
    hipLaunchKernelGGL(calib_encode_kernel<false>, grid, dim3(256), 0, s, enc, stride, calib,
                       n_cal, amax);
  hip_check(hipGetLastError(), "calib_encode_kernel launch");
}

void launch_calibrate_amax(const RelightArgs &a, hipStream_t s) {
  require(a.n_pixels >= 1 && a.k >= 1 && a.k <= NWPRT_MAX_K && a.features && a.mk &&
              a.w_tiles && a.amax && (!a.aux || a.waux_tiles),
          NWPRT_ERR_INVALID_ARGUMENT, "calibration: missing operand");
  constexpr int kChunk = 16;
  const dim3 grid((a.n_pixels + kRelightBlock - 1) / kRelightBlock, (a.k + kChunk - 1) / kChunk);
  if (a.aux)
    hipLaunchKernelGGL(calib_amax_kernel<true>, grid, dim3(kRelightBlock), 0, s, a, kChunk);
  else
    hipLaunchKernelGGL(calib_amax_kernel<false>, grid, dim3(kRelightBlock), 0, s, a, kChunk);
  hip_check(hipGetLastError(), "calib_amax_kernel launch");
}

void launch_tonemap(const float *src, uint32_t *dst, int n_pixels, hipStream_t s) {
  require(n_pixels >= 1, NWPRT_ERR_INVALID_ARGUMENT, "tonemap: no pixels");
  hipLaunchKernelGGL(tonemap_kernel, dim3((n_pixels + 255) / 256), dim3(256), 0, s, src, dst,
                     n_pixels);
  hip_check(hipGetLastError(), "tonemap_kernel launch");
}

}  // namespace nwprt
