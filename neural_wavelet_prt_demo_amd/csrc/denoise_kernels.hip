// Guided edge-avoiding a-trous denoiser: the replacement for the denoiser
// step after the relight pass (Renderer/RenderEngine.cpp:390-403 calling
// OptixBackend::denoise, Renderer/OptixBackend.cpp:841-906).  The
// reference's denoiser is a closed learned model; this one is the published
// edge-avoiding a-trous wavelet filter (Dammertz, Sewtz, Hanika, Lensch,
// "Edge-Avoiding A-Trous Wavelet Transform for fast Global Illumination
// Filtering", HPG 2010) guided by the same inputs the reference passes:
// the colour, the albedo and the normal buffers (float4, guideAlbedo and
// guideNormal on, OptixBackend.cpp:783-785), with the alpha channel copied
// from the input (OPTIX_DENOISER_ALPHA_MODE_COPY, OptixBackend.cpp:815) and
// the result blended with the input by blendFactor.
//
// The filter, for an image of P = W x H pixels:
// - HDR normalisation: kappa = 2^-m with m the mean of log2(Y) over pixels
//   whose luminance Y = 0.2126 R + 0.7152 G + 0.0722 B is > 1e-8 (kappa = 1
//   if none), so colour distances are relative to the image's log-average
//   luminance;
// - passes i = 0 .. L-1 with step s = 2^i over the 5 x 5 B3-spline taps
//   h(dx) h(dy), h = (1, 4, 6, 4, 1) / 16, skipping taps outside the image:
//     w(p, q) = h exp(-(kappa^2 |c_q - c_p|^2 4^i / sc^2
//                      + |n_q - n_p|^2 / sn^2 + |a_q - a_p|^2 / sa^2)),
//     c'_p = sum_q w c_q / sum_q w   (RGB; c the previous pass's output),
//   the colour sigma halving each pass as in the paper;
// - output RGB = blend * input + (1 - blend) * c after pass L-1, alpha =
//   the input's alpha.
// Everything runs in FP32: one exp2 per tap in the passes, and the
// normalisation's sum of log2(Y) is compensated (each partial carries the
// rounding error of its additions, Knuth's TwoSum), with exact integer
// counts, reduced in a fixed order.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace nwprt {
namespace {

constexpr int kTile = 32;            // 32 x 32 pixels per workgroup
constexpr int kIntensityBlocks = 512;
constexpr int kIntensityThreads = 256;

// A compensated FP32 sum: hi + lo carries the exact sum of what was added up
// to the rounding of lo, which only collects the TwoSum errors.
struct CSum {
  float hi, lo;
};

__device__ __forceinline__ CSum csum_add(CSum a, float b) {
  const float s = a.hi + b, bb = s - a.hi;
  const float e = (a.hi - (s - bb)) + (b - bb);
  return {s, a.lo + e};
}

__device__ __forceinline__ CSum csum_merge(CSum a, CSum b) {
  CSum r = csum_add(a, b.hi);
  r.lo += b.lo;
  return r;
}

// Per-block partial sums of log2(Y) and counts over lit pixels.
__global__ __launch_bounds__(kIntensityThreads) void intensity_partials_kernel(
    const float4 *__restrict__ color, long long n, float4 *__restrict__ partials) {
  CSum s = {0.0f, 0.0f};
  unsigned c = 0;
  for (long long p = (long long)blockIdx.x * kIntensityThreads + threadIdx.x; p < n;
       p += (long long)kIntensityBlocks * kIntensityThreads) {
    const float4 v = color[p];
    const float y = fmaf(0.0722f, v.z, fmaf(0.7152f, v.y, 0.2126f * v.x));
    if (y > 1e-8f) {
      s = csum_add(s, log2f(y));
      ++c;
    }
  }
  __shared__ CSum ss[kIntensityThreads];
  __shared__ unsigned sc[kIntensityThreads];
  ss[threadIdx.x] = s;
  sc[threadIdx.x] = c;
  __syncthreads();
  for (int h = kIntensityThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      ss[threadIdx.x] = csum_merge(ss[threadIdx.x], ss[threadIdx.x + h]);
      sc[threadIdx.x] += sc[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    partials[blockIdx.x] = make_float4(ss[0].hi, ss[0].lo, __uint_as_float(sc[0]), 0.0f);
}

// kappa = 2^-(sum / count) from the partials, in a fixed order.  Counts stay
// below 2^32 (images are limited to 2^31 pixels).
__global__ __launch_bounds__(kIntensityBlocks) void intensity_finish_kernel(
    const float4 *__restrict__ partials, float *__restrict__ kappa) {
  __shared__ CSum ss[kIntensityBlocks];
  __shared__ unsigned sc[kIntensityBlocks];
  const float4 v = partials[threadIdx.x];
  ss[threadIdx.x] = {v.x, v.y};
  sc[threadIdx.x] = __float_as_uint(v.z);
  __syncthreads();
  for (int h = kIntensityBlocks / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      ss[threadIdx.x] = csum_merge(ss[threadIdx.x], ss[threadIdx.x + h]);
      sc[threadIdx.x] += sc[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    kappa[0] = sc[0] > 0 ? exp2f(-((ss[0].hi + ss[0].lo) / (float)sc[0])) : 1.0f;
}

__device__ __forceinline__ float sq3(float4 a, float4 b) {
  const float x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
  return x * x + y * y + z * z;
}

// One a-trous pass on an LDS tile.  With step s the taps of a pixel stay on
// its residue class (x mod s, y mod s), where they are a dense 5 x 5
// neighbourhood: workgroup (tx * s + rx, ty * s + ry) owns the T x T
// lattice points (tx * T + i, ty * T + j) of residue (rx, ry), T = kTile,
// i.e. pixels (rx + (tx * T + i) s, ry + (ty * T + j) s), and first loads
// the (T + 4) x (T + 4) halo of colour, normal and albedo into LDS (61 KB
// at T = 32; 32 x 32 tiles re-read 1.27x of the guides against 1.56x at
// 16 x 16, and ran the five 1080p passes in 0.33 instead of 0.43 ms).  Workgroups of the
// same tile and neighbouring residues are adjacent in the grid, so they
// share cache lines in L2.  The grid is 1-D, gx x gy logical workgroups
// padded to a multiple of 8: the hardware hands workgroup b to XCD b % 8,
// and logical workgroup (b % 8) * per + b / 8 keeps each XCD on one
// contiguous range of tiles and residues (one L2 per XCD).  inv_c2 =
// 4^i / sc^2 (times kappa^2 on the device), inv_n2 = 1 / sn^2, inv_a2 =
// 1 / sa^2.  LAST: blend with `orig` and write `out`; otherwise write the
// pass result (alpha from `orig`).
template <bool LAST>
__global__ __launch_bounds__(kTile * kTile) void atrous_kernel(
    const float4 *__restrict__ in, const float4 *__restrict__ orig,
    const float4 *__restrict__ normal, const float4 *__restrict__ albedo,
    const float *__restrict__ kappa, float4 *__restrict__ out, int w, int h, int step,
    int gx, int gy, float inv_c2, float inv_n2, float inv_a2, float blend) {
  constexpr int kHalo = kTile + 4;
  __shared__ float4 sc[kHalo * kHalo], sn[kHalo * kHalo], sa[kHalo * kHalo];
  const int per = (int)gridDim.x / 8;
  const int lb = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (lb >= gx * gy) return;  // padding workgroup
  const int bx = lb % gx, by = lb / gx;
  const int rx = bx % step, tx = bx / step;
  const int ry = by % step, ty = by / step;
  const int x0 = rx + tx * kTile * step, y0 = ry + ty * kTile * step;  // lattice (0, 0)
  if (x0 >= w || y0 >= h) return;  // whole workgroup outside: no barrier reached
  const int tid = threadIdx.y * kTile + threadIdx.x;
  for (int i = tid; i < kHalo * kHalo; i += kTile * kTile) {
    const int qx = x0 + (i % kHalo - 2) * step, qy = y0 + (i / kHalo - 2) * step;
    if (qx >= 0 && qx < w && qy >= 0 && qy < h) {
      const size_t q = (size_t)qy * w + qx;
      sc[i] = in[q];
      sn[i] = normal[q];
      sa[i] = albedo[q];
    }
  }
  __syncthreads();
  const int x = x0 + threadIdx.x * step, y = y0 + threadIdx.y * step;
  if (x >= w || y >= h) return;
  const float k = kappa[0];
  const float wc = k * k * inv_c2 * 1.44269504088896341f;  // exp(-d) = exp2(-d log2 e)
  const float wn = inv_n2 * 1.44269504088896341f, wa = inv_a2 * 1.44269504088896341f;
  const int c0 = (threadIdx.y + 2) * kHalo + threadIdx.x + 2;
  const float4 cp = sc[c0], np = sn[c0], ap = sa[c0];
  constexpr float kh[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
  float sr = 0.f, sg = 0.f, sb = 0.f, sw = 0.f;
#pragma unroll
  for (int dy = -2; dy <= 2; ++dy) {
    const int qy = y + dy * step;
    if (qy < 0 || qy >= h) continue;
#pragma unroll
    for (int dx = -2; dx <= 2; ++dx) {
      const int qx = x + dx * step;
      if (qx < 0 || qx >= w) continue;
      const int q = c0 + dy * kHalo + dx;
      const float4 cq = sc[q], nq = sn[q], aq = sa[q];
      const float d = wc * sq3(cq, cp) + wn * sq3(nq, np) + wa * sq3(aq, ap);
      const float wt = kh[dx + 2] * kh[dy + 2] * __builtin_amdgcn_exp2f(-d);
      sr += wt * cq.x;
      sg += wt * cq.y;
      sb += wt * cq.z;
      sw += wt;
    }
  }
  // The centre tap has weight h0^2 > 0, so sw > 0.
  const float inv = 1.0f / sw;
  const size_t p = (size_t)y * w + x;
  const float4 o = orig[p];
  float4 r;
  if (LAST) {
    const float b1 = 1.0f - blend;
    r = make_float4(blend * o.x + b1 * (sr * inv), blend * o.y + b1 * (sg * inv),
                    blend * o.z + b1 * (sb * inv), o.w);
  } else {
    r = make_float4(sr * inv, sg * inv, sb * inv, o.w);
  }
  out[p] = r;
}

}  // namespace

size_t denoise_scratch_bytes(int w, int h) {
  return 2 * (size_t)w * h * sizeof(float4) + kIntensityBlocks * sizeof(float4) + 256;
}

void launch_denoise(const DenoiseArgs &a, void *scratch, hipStream_t s) {
  require(a.color && a.normal && a.albedo && a.out && scratch, NWPRT_ERR_INVALID_ARGUMENT,
          "denoise: null buffer");
  require(a.width > 0 && a.height > 0 && a.iterations >= 1 && a.iterations <= kMaxDenoisePasses,
          NWPRT_ERR_INVALID_ARGUMENT, "denoise: bad size or iteration count");
  const size_t P = (size_t)a.width * a.height;
  float4 *t0 = (float4 *)scratch, *t1 = t0 + P;
  float4 *partials = (float4 *)(t1 + P);
  float *kappa = (float *)(partials + kIntensityBlocks);
  hipLaunchKernelGGL(intensity_partials_kernel, dim3(kIntensityBlocks), dim3(kIntensityThreads),
                     0, s, (const float4 *)a.color, (long long)P, partials);
  hip_check(hipGetLastError(), "intensity_partials_kernel launch");
  hipLaunchKernelGGL(intensity_finish_kernel, dim3(1), dim3(kIntensityBlocks), 0, s,
                     (const float4 *)partials, kappa);
  hip_check(hipGetLastError(), "intensity_finish_kernel launch");
  const dim3 block(kTile, kTile);
  const float4 *orig = (const float4 *)a.color;
  const float4 *src = orig;
  const float inv_n2 = 1.0f / (a.sigma_normal * a.sigma_normal);
  const float inv_a2 = 1.0f / (a.sigma_albedo * a.sigma_albedo);
  for (int i = 0; i < a.iterations; ++i) {
    const float inv_c2 = (float)(1 << (2 * i)) / (a.sigma_color * a.sigma_color);
    const bool last = i == a.iterations - 1;
    // Per residue class: ceil(ceil(W / s) / 16) x ceil(ceil(H / s) / 16) tiles.
    const int st = 1 << i;
    const int lw = (a.width + st - 1) / st, lh = (a.height + st - 1) / st;
    const int gx = ((lw + kTile - 1) / kTile) * st, gy = ((lh + kTile - 1) / kTile) * st;
    const dim3 grid((unsigned)((((long long)gx * gy + 7) / 8) * 8));
    float4 *dst = last ? (float4 *)a.out : (i % 2 == 0 ? t0 : t1);
    if (last)
      hipLaunchKernelGGL(atrous_kernel<true>, grid, block, 0, s, src, orig,
                         (const float4 *)a.normal, (const float4 *)a.albedo, kappa, dst, a.width,
                         a.height, st, gx, gy, inv_c2, inv_n2, inv_a2, a.blend);
    else
      hipLaunchKernelGGL(atrous_kernel<false>, grid, block, 0, s, src, orig,
                         (const float4 *)a.normal, (const float4 *)a.albedo, kappa, dst, a.width,
                         a.height, st, gx, gy, inv_c2, inv_n2, inv_a2, a.blend);
    hip_check(hipGetLastError(), "atrous_kernel launch");
    src = dst;
  }
}

}  // namespace nwprt
