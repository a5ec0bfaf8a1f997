// Per-pixel encoding for gfx950: multiresolution hash-grid features of the
// first-hit position, then the optional 32 x 32 projection.  One lane per
// pixel, computed once per pixel (the reference encodes every pixel K times,
// once per replicated (pixel, wavelet) column: Renderer/tinycudann/
// utils_kernels.h:25 and prt_network.h:86-91).
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hashgrid_device.hpp"
#include "kernels.hpp"

namespace nwprt {
namespace {

constexpr int kEncodeBlock = 256;

// Row p of a [n][W] array in element type OUT (W a multiple of 8).
template <Elem OUT, int W>
__device__ __forceinline__ void store_row(void *base, int p, const float *v) {
  if (OUT == Elem::F32) {
    float4 *dst = (float4 *)((float *)base + (size_t)p * W);
#pragma unroll
    for (int q = 0; q < W / 4; ++q) dst[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
  } else {
    dev::u32x4 *dst = (dev::u32x4 *)((uint16_t *)base + (size_t)p * W);
#pragma unroll
    for (int q = 0; q < W / 8; ++q) {
      dev::u32x4 r;
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const float lo = v[8 * q + 2 * d], hi = v[8 * q + 2 * d + 1];
        r[d] = OUT == Elem::BF16 ? dev::pack_bf16(lo, hi) : dev::pack_f16_rounded(lo, hi);
      }
      dst[q] = r;
    }
  }
}

// Planar form (16 levels x 2 features): workgroup b encodes levels q and
// 15 - q (q = b % 8) of pixels [256 (b / 8), +256) and writes their four
// raw features to plane q.  Workgroups are dealt to the 8 XCDs round-robin,
// so every workgroup of plane q runs on XCD q and that XCD's L2 only ever
// holds these two levels' entries (at most 2 x 2^19 x 2 x 2 B = 4 MiB for
// FP16 hashed levels).  Pairing a fine level with a coarse one evens the
// work per plane.
template <bool HF16, Elem OUT>
__global__ __launch_bounds__(kEncodeBlock) void encode_planar_kernel(EncodeArgs a) {
  const int q = blockIdx.x & 7;
  const int p = (blockIdx.x >> 3) * kEncodeBlock + threadIdx.x;
  if (p >= a.n_pixels) return;
  const bool hit = a.env_flag == nullptr || a.env_flag[p] != 0;
  float v[4] = {0.f, 0.f, 0.f, 0.f};
  if (hit) {
    const float4 x4 = *(const float4 *)(a.position + 4 * (size_t)p);
    const float x[3] = {x4.x, x4.y, x4.z};
    dev::accumulate_level<2, HF16>(a.lv, a.hash, q, x, v, nullptr);
    dev::accumulate_level<2, HF16>(a.lv, a.hash, 15 - q, x, v + 2, nullptr);
  }
  const size_t at = (size_t)q * a.n_pixels + p;
  if (OUT == Elem::BF16)
    ((uint2 *)a.features)[at] = make_uint2(dev::pack_bf16(v[0], v[1]), dev::pack_bf16(v[2], v[3]));
  else
    ((uint2 *)a.features)[at] =
        make_uint2(dev::pack_f16_rounded(v[0], v[1]), dev::pack_f16_rounded(v[2], v[3]));
}

template <int F, bool HF16, Elem OUT, bool DBG>
__global__ __launch_bounds__(kEncodeBlock) void encode_kernel(EncodeArgs a) {
  constexpr int L = 32 / F;
  __shared__ float proj[32 * 32];
  const bool has_proj = a.projection != nullptr;
  if (has_proj)
    for (int i = threadIdx.x; i < 32 * 32; i += kEncodeBlock) proj[i] = a.projection[i];
  __syncthreads();

  const int p = dev::xcd_tile(blockIdx.x, gridDim.x) * kEncodeBlock + threadIdx.x;
  if (p >= a.n_pixels) return;
  const bool hit = a.env_flag == nullptr || a.env_flag[p] != 0;

  float feat[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) feat[i] = 0.0f;

  if (hit || DBG) {
    const float4 x4 = *(const float4 *)(a.position + 4 * (size_t)p);
    const float x[3] = {x4.x, x4.y, x4.z};
#pragma unroll
    for (int l = 0; l < L; ++l)
      dev::accumulate_level<F, HF16>(a.lv, a.hash, l, x, feat + l * F,
                                DBG ? a.dbg_indices + ((size_t)p * L + l) * 8 : nullptr);
    if (DBG && a.dbg_features) {
#pragma unroll
      for (int i = 0; i < 32; ++i) a.dbg_features[(size_t)p * 32 + i] = feat[i];
    }
  }
  if (!hit) {
#pragma unroll
    for (int i = 0; i < 32; ++i) feat[i] = 0.0f;
  }

  float out[32];
  if (has_proj) {
#pragma unroll
    for (int o = 0; o < 32; ++o) {
      float acc = 0.0f;
#pragma unroll
      for (int i = 0; i < 32; ++i) acc = fmaf(proj[o * 32 + i], feat[i], acc);
      out[o] = acc;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 32; ++i) out[i] = feat[i];
  }

  store_row<OUT, 32>(a.features, p, out);

  if (a.aux) {
    // The reference's per-pair input rows 3..15 (Renderer/tinycudann/
    // utils_kernels.h:31-46, INPUT_BUFFER_* of prt_globals.h): normal and
    // reflection as stored, log(25 r + 1) / log 26, 2 kd - 1, 2 ks - 1; then
    // three padding inputs.  Computed once per pixel.
    const float4 nn = *(const float4 *)(a.normal + 4 * (size_t)p);
    const float4 rr = *(const float4 *)(a.reflect + 4 * (size_t)p);
    const float4 kd = *(const float4 *)(a.kd + 4 * (size_t)p);
    const float4 ks = *(const float4 *)(a.ks + 4 * (size_t)p);
    const float rough = a.roughness[p];
    const float rgh_den = 1.0f / logf(26.0f);
    const float v[16] = {nn.x, nn.y, nn.z, rr.x, rr.y, rr.z,
                         logf(fmaf(25.0f, rough, 1.0f)) * rgh_den,
                         fmaf(2.0f, kd.x, -1.0f), fmaf(2.0f, kd.y, -1.0f), fmaf(2.0f, kd.z, -1.0f),
                         fmaf(2.0f, ks.x, -1.0f), fmaf(2.0f, ks.y, -1.0f), fmaf(2.0f, ks.z, -1.0f),
                         a.aux_pad, a.aux_pad, a.aux_pad};
    store_row<OUT, 16>(a.aux, p, v);
  }
}

template <int F, bool HF16, bool DBG>
void dispatch_out(const EncodeArgs &a, dim3 grid, hipStream_t s) {
  switch (a.out_elem) {
    case Elem::F32:
      hipLaunchKernelGGL((encode_kernel<F, HF16, Elem::F32, DBG>), grid, dim3(kEncodeBlock), 0, s, a);
      break;
    case Elem::F16:
      hipLaunchKernelGGL((encode_kernel<F, HF16, Elem::F16, DBG>), grid, dim3(kEncodeBlock), 0, s, a);
      break;
    case Elem::BF16:
      hipLaunchKernelGGL((encode_kernel<F, HF16, Elem::BF16, DBG>), grid, dim3(kEncodeBlock), 0, s, a);
      break;
  }
}

template <int F>
void dispatch_f(const EncodeArgs &a, dim3 grid, hipStream_t s) {
  const bool dbg = a.dbg_indices != nullptr;
  if (a.hash_f16) {
    if (dbg) dispatch_out<F, true, true>(a, grid, s);
    else dispatch_out<F, true, false>(a, grid, s);
  } else {
    if (dbg) dispatch_out<F, false, true>(a, grid, s);
    else dispatch_out<F, false, false>(a, grid, s);
  }
}

}  // namespace

void launch_encode(const EncodeArgs &a, hipStream_t s) {
  require(a.n_pixels >= 1, NWPRT_ERR_INVALID_ARGUMENT, "encode: no pixels");
  require(a.lv.n_levels * a.n_features == 32, NWPRT_ERR_UNSUPPORTED,
          "encode: n_levels * n_features must be 32");
  for (int l = 0; l < a.lv.n_levels; ++l)
    require(a.lv.dense[l] || (a.lv.size[l] & (a.lv.size[l] - 1)) == 0, NWPRT_ERR_INVALID_ARGUMENT,
            "encode: hashed level size must be a power of two");
  const dim3 grid((a.n_pixels + kEncodeBlock - 1) / kEncodeBlock);
  if (a.planar) {
    require(a.n_features == 2 && a.aux == nullptr && a.dbg_indices == nullptr &&
                (a.out_elem == Elem::BF16 || a.out_elem == Elem::F16),
            NWPRT_ERR_INVALID_ARGUMENT,
            "encode: the planar form is for 16 x 2 grids, BF16/FP16, no aux or tap");
    require((long long)grid.x * 8 < (1ll << 31), NWPRT_ERR_INVALID_ARGUMENT, "encode: too many pixels");
    const dim3 g8(grid.x * 8);
    if (a.hash_f16) {
      if (a.out_elem == Elem::BF16)
        hipLaunchKernelGGL((encode_planar_kernel<true, Elem::BF16>), g8, dim3(kEncodeBlock), 0, s, a);
      else
        hipLaunchKernelGGL((encode_planar_kernel<true, Elem::F16>), g8, dim3(kEncodeBlock), 0, s, a);
    } else {
      if (a.out_elem == Elem::BF16)
        hipLaunchKernelGGL((encode_planar_kernel<false, Elem::BF16>), g8, dim3(kEncodeBlock), 0, s, a);
      else
        hipLaunchKernelGGL((encode_planar_kernel<false, Elem::F16>), g8, dim3(kEncodeBlock), 0, s, a);
    }
    hip_check(hipGetLastError(), "encode_planar_kernel launch");
    return;
  }
  switch (a.n_features) {
    case 1: dispatch_f<1>(a, grid, s); break;
    case 2: dispatch_f<2>(a, grid, s); break;
    case 4: dispatch_f<4>(a, grid, s); break;
    case 8: dispatch_f<8>(a, grid, s); break;
    default: fail(NWPRT_ERR_UNSUPPORTED, "encode: n_features must be 1, 2, 4 or 8");
  }
  hip_check(hipGetLastError(), "encode_kernel launch");
}

}  // namespace nwprt
