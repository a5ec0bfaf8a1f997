// Environment-map kernels for gfx950:
//  - non-standard 2-D Haar transform of each cubemap face (FP64 scaling
//    coefficients, one rounding to FP32 per output coefficient);
//  - area-weighted top-K selection with ties broken by lower flat index;
//  - per-relight preparation of the selected wavelets' feature rows.
//
// The reference only loads the results of an offline transform
// (waveletcoeffs/, waveletstrength/: Renderer/EnvironmentMap.cpp:36-40, 56-67)
// and gathers feature rows per pair (Renderer/tinycudann/utils_kernels.h:97-115).
#include <hip/hip_runtime.h>

#include <cmath>

#include "device_common.hpp"
#include "kernels.hpp"

// No floating-point contraction in this file: the Haar scaling sums, the
// top-K scores and the rotation's direction arithmetic are each defined as
// separately rounded operations, the way the oracle evaluates them, so that
// index sets and texel choices agree bit for bit.
#pragma clang fp contract(off)

namespace nwprt {

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    fail(NWPRT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

// One 2x2 step of the orthonormal non-standard Haar transform.  a = x[2i][2j],
// b = x[2i][2j+1], c = x[2i+1][2j], d = x[2i+1][2j+1] (row 0 = bottom row of
// the face).  The row step then the column step, each with 1/sqrt(2), equal
// one division by 2; the operation order below is fixed so that any FP64
// implementation following it produces identical bits.
struct Haar4 {
  double s, h, v, dg;
};
__device__ __forceinline__ Haar4 haar2x2(double a, double b, double c, double d) {
  Haar4 r;
  r.s = ((a + b) + (c + d)) * 0.5;
  r.h = ((a - b) + (c - d)) * 0.5;
  r.v = ((a + b) - (c + d)) * 0.5;
  r.dg = ((a - b) - (c - d)) * 0.5;
  return r;
}

// Mallat layout: at a level whose LL block is s x s (hs = s / 2), the details
// of 2x2 block (i, j) go to (i, hs + j) [h], (hs + i, j) [v] and
// (hs + i, hs + j) [dg]; the final LL is the DC coefficient at (0, 0) = face
// sum / N.  Each output is rounded to FP32 once.
__device__ __forceinline__ void put_details(float *dst, int n, int hs, int i, int j,
                                            const Haar4 &r) {
  dst[((size_t)i * n + hs + j) * 3] = (float)r.h;
  dst[((size_t)(hs + i) * n + j) * 3] = (float)r.v;
  dst[((size_t)(hs + i) * n + hs + j) * 3] = (float)r.dg;
}

// The transform in two kernels.  A face splits into tiles of T x T texels
// (T = min(N, 32)); the levels inside a tile need only its own texels.
// haar_tile_kernel: one workgroup per (map, face, tile), all three channels:
//   level 1 reads the tile's FP32 texels (one 2x2 block and its three
//   channels per thread, rows read as contiguous 24-byte pairs), then the LL
//   values (FP64, in LDS) go down to one per tile and channel, which goes to
//   `ll_out` [map][face][tile][3] (FP64).
// haar_top_kernel: one workgroup per (map, face) runs the levels above the
//   tiles on that (N/T) x (N/T) LL block and writes the DC.
// The per-block arithmetic is haar2x2 on the same FP64 values whatever the
// tiling, so the coefficients are those of a whole-face transform, bit for
// bit.  At N = 128 that is 96 workgroups per map instead of one per (face,
// channel).
constexpr int kHaarTile = 32;
constexpr int kHaarTileBlock = (kHaarTile / 2) * (kHaarTile / 2);  // 256

__global__ __launch_bounds__(kHaarTileBlock) void haar_tile_kernel(const float *__restrict__ faces,
                                                                   size_t map_stride,
                                                                   float *__restrict__ coef,
                                                                   int n,
                                                                   double *__restrict__ ll_out) {
  __shared__ double ll[3][kHaarTileBlock];
  const int T = n < kHaarTile ? n : kHaarTile, tpf = n / T, h1 = T / 2;
  const int tile = blockIdx.x, face = blockIdx.y, map = blockIdx.z;
  const int ty = tile / tpf, tx = tile % tpf;
  const size_t base = (size_t)map * map_stride + (size_t)face * n * n * 3;
  const float *src = faces + base;
  float *dst = coef + base;
  const int tid = threadIdx.x;
  if (tid < h1 * h1) {
    const int i = tid / h1, j = tid % h1;
    const int gi = ty * h1 + i, gj = tx * h1 + j;
    const float *r0 = src + ((size_t)(2 * gi) * n + 2 * gj) * 3;
    const float *r1 = r0 + (size_t)n * 3;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const Haar4 r = haar2x2(r0[ch], r0[3 + ch], r1[ch], r1[3 + ch]);
      ll[ch][i * h1 + j] = r.s;
      put_details(dst + ch, n, n / 2, gi, gj, r);
    }
  }
  __syncthreads();
  // Local LL size ls -> ls / 2; the global LL at this level is tpf * ls.
  for (int ls = h1; ls >= 2; ls >>= 1) {
    const int hl = ls >> 1, hs = tpf * hl;
    const bool act = tid < 3 * hl * hl;
    const int ch = act ? tid / (hl * hl) : 0, q = tid % (hl * hl);
    const int i = q / hl, j = q % hl;
    double keep = 0.0;
    if (act) {
      const double *L = ll[ch];
      const Haar4 r = haar2x2(L[(2 * i) * h1 + 2 * j], L[(2 * i) * h1 + 2 * j + 1],
                              L[(2 * i + 1) * h1 + 2 * j], L[(2 * i + 1) * h1 + 2 * j + 1]);
      keep = r.s;
      put_details(dst + ch, n, hs, ty * hl + i, tx * hl + j, r);
    }
    __syncthreads();  // every read of this level is done before LL is overwritten
    if (act) ll[ch][i * h1 + j] = keep;
    __syncthreads();
  }
  if (tid < 3)
    ll_out[(((size_t)map * 6 + face) * tpf * tpf + tile) * 3 + tid] = ll[tid][0];
}

__global__ __launch_bounds__(64) void haar_top_kernel(const double *__restrict__ ll_in,
                                                      size_t map_stride,
                                                      float *__restrict__ coef, int n) {
  constexpr int kMaxTpf = NWPRT_MAX_FACE_SIZE / kHaarTile;  // 4
  __shared__ double ll[3][kMaxTpf * kMaxTpf];
  const int T = n < kHaarTile ? n : kHaarTile, tpf = n / T;
  const int face = blockIdx.x, map = blockIdx.y;
  float *dst = coef + (size_t)map * map_stride + (size_t)face * n * n * 3;
  const int tid = threadIdx.x;
  const double *src = ll_in + ((size_t)map * 6 + face) * tpf * tpf * 3;
  if (tid < 3 * tpf * tpf) ll[tid % 3][tid / 3] = src[tid];
  __syncthreads();
  for (int ls = tpf; ls >= 2; ls >>= 1) {
    const int hs = ls >> 1;
    const bool act = tid < 3 * hs * hs;
    const int ch = act ? tid / (hs * hs) : 0, q = tid % (hs * hs);
    const int i = q / hs, j = q % hs;
    double keep = 0.0;
    if (act) {
      const double *L = ll[ch];
      const Haar4 r = haar2x2(L[(2 * i) * tpf + 2 * j], L[(2 * i) * tpf + 2 * j + 1],
                              L[(2 * i + 1) * tpf + 2 * j], L[(2 * i + 1) * tpf + 2 * j + 1]);
      keep = r.s;
      put_details(dst + ch, n, hs, i, j, r);
    }
    __syncthreads();
    if (act) ll[ch][i * tpf + j] = keep;
    __syncthreads();
  }
  if (tid < 3) dst[tid] = (float)ll[tid][0];
}

// ------------------------------------------------------------------ top-K --

constexpr int kTopkBlock = 1024;

// Score of coefficient t: |luminance| (weights of Renderer/EnvironmentMap.cpp:149-151,
// evaluated unfused in FP32 as there) times the support area of its basis
// function relative to the face, 4^-j for a detail at Mallat position with
// max(row, col) in [2^j, 2^(j+1)) and 1 for the DC term.  Scores below the
// smallest normal float and NaNs become 0.
// n is a power of two (face sizes are), so the position math is shifts.
__device__ __forceinline__ uint32_t score_bits(const float *__restrict__ coef, int t, int n) {
  const float r = coef[3 * (size_t)t], g = coef[3 * (size_t)t + 1], b = coef[3 * (size_t)t + 2];
  const float lum = r * 0.212671f + g * 0.715160f + b * 0.072169f;
  const int log2n = __ffs(n) - 1;
  const int q = t & (n * n - 1);
  const int row = q >> log2n, col = q & (n - 1);
  const int m = row > col ? row : col;
  const int j = m == 0 ? 0 : 31 - __clz(m);
  float s = ldexpf(fabsf(lum), -2 * j);
  if (!(s >= 1.17549435e-38f)) s = 0.0f;  // NaN, 0 and subnormal scores
  return __float_as_uint(s);
}

// Key: score bits in the high word, (2^32 - 1 - index) in the low word, so a
// larger key means a larger score, and for equal scores a lower index.  Keys
// are unique, so "the K largest keys" is an exact set.
__device__ __forceinline__ uint64_t topk_key(const uint32_t *scores, int t) {
  return ((uint64_t)scores[t] << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)t);
}

// Bucket of the digit at `shift` that holds the need-th largest remaining
// key (need = *s_need on entry): with suf[d] = sum of hist[d..255], the one
// bucket d with suf[d] >= need > suf[d + 1] extends the prefix, and *s_done
// is set when that whole bucket is taken.  Thread t < 256 takes bucket
// 255 - t, so suffix sums over buckets are prefix sums over threads: a
// shuffle scan within each wave, then the totals of the waves before it
// (`wt`, 4 words of LDS).  Every thread of the workgroup calls it, after a
// barrier that made `hist` visible.
__device__ void select_digit(const uint32_t *hist, uint32_t *wt, int shift, uint64_t *s_prefix,
                             uint64_t *s_mask, uint32_t *s_need, int *s_done) {
  const int tid = threadIdx.x, lane = tid & 63;
  const uint32_t need = *s_need;
  uint32_t v = 0, p = 0;
  if (tid < 256) {
    v = hist[255 - tid];
    p = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t u = __shfl_up(p, off, 64);
      if (lane >= off) p += u;
    }
    if (lane == 63) wt[tid >> 6] = p;
  }
  __syncthreads();
  if (tid < 256) {
    for (int w = 0; w < (tid >> 6); ++w) p += wt[w];
    // p = suf[255 - tid]; p - v = suf[256 - tid] (0 for the top bucket).
    if (p >= need && p - v < need) {
      const uint32_t d = 255u - (uint32_t)tid;
      *s_need = need - (p - v);
      *s_prefix |= (uint64_t)d << shift;
      *s_mask |= 0xFFull << shift;
      *s_done = v == need - (p - v);
    }
  }
  __syncthreads();
}

// One workgroup: MSB-first radix select of the K-th largest key (8-bit
// digits, early exit once a whole bucket is taken) -> compaction -> rank
// sort of the K selected keys in LDS (descending).
// - Scores and the first digit's per-workgroup histograms come from
//   score_kernel; this kernel sums the histograms.
// - Keys above the chosen first-digit bucket are taken at once; keys in it
//   (at most kTopkCand of them) move to LDS, and the remaining digits are
//   resolved there.  A larger bucket falls back to passes over all keys.
constexpr int kTopkCand = 3072;
constexpr int kScoreBlock = 256;

// Scores of all 6 n^2 coefficients across the machine, kScoreKeys per
// workgroup, and each workgroup's histogram of the keys' first digit (the
// top byte of the score) written to wg_hist[blockIdx.x][256] -- per wave in
// LDS first, since float exponents bunch keys into a few buckets.  The
// selection below then starts from these histograms on one workgroup.
__global__ __launch_bounds__(kScoreBlock) void score_kernel(const float *__restrict__ coef, int n,
                                                            uint32_t *__restrict__ scores,
                                                            uint32_t *__restrict__ wg_hist) {
  __shared__ uint32_t wh[kScoreBlock / 64][256];
  const int tid = threadIdx.x, wave = tid >> 6;
  for (int i = tid; i < (kScoreBlock / 64) * 256; i += kScoreBlock) (&wh[0][0])[i] = 0u;
  __syncthreads();
  const int M = 6 * n * n;
  const int t0 = blockIdx.x * kScoreKeys;
  for (int t = t0 + tid; t < t0 + kScoreKeys && t < M; t += kScoreBlock) {
    const uint32_t sc = score_bits(coef, t, n);
    scores[t] = sc;
    atomicAdd(&wh[wave][sc >> 24], 1u);
  }
  __syncthreads();
  uint32_t h = 0;
#pragma unroll
  for (int w = 0; w < kScoreBlock / 64; ++w) h += wh[w][tid];
  wg_hist[(size_t)blockIdx.x * 256 + tid] = h;
}

__global__ __launch_bounds__(kTopkBlock) void topk_kernel(const float *__restrict__ coef, int n,
                                                          int K, uint32_t *__restrict__ scores,
                                                          const uint32_t *__restrict__ wg_hist,
                                                          int n_wg,
                                                          int32_t *__restrict__ out_idx,
                                                          float *__restrict__ out_str) {
  __shared__ uint32_t hist[256], wt[4];
  __shared__ uint64_t sel[NWPRT_MAX_K], sorted[NWPRT_MAX_K];
  __shared__ uint64_t cand[kTopkCand];
  __shared__ uint64_t s_prefix, s_mask;
  __shared__ uint32_t s_need, s_count, s_ncand;
  __shared__ int s_done;
  const int tid = threadIdx.x;
  const int M = 6 * n * n;

  if (tid == 0) {
    s_prefix = 0;
    s_mask = 0;
    s_need = (uint32_t)K;
    s_done = 0;
    s_count = 0;
    s_ncand = 0;
  }
  if (tid < 256) {
    uint32_t h = 0;
    for (int w = 0; w < n_wg; ++w) h += wg_hist[(size_t)w * 256 + tid];
    hist[tid] = h;
  }
  __syncthreads();
  select_digit(hist, wt, 56, &s_prefix, &s_mask, &s_need, &s_done);
  const uint32_t d1 = (uint32_t)(s_prefix >> 56);
  const bool in_lds = hist[d1] <= (uint32_t)kTopkCand;
  __syncthreads();  // every thread has read hist before it is reused

  // Remaining digits over `n_keys` keys from `key_at`.
  auto resolve = [&](int n_keys, auto key_at) {
    for (int pass = 1; pass < 8 && !s_done; ++pass) {
      const int shift = 56 - 8 * pass;
      if (tid < 256) hist[tid] = 0;
      __syncthreads();
      const uint64_t prefix = s_prefix, mask = s_mask;
      for (int i = tid; i < n_keys; i += kTopkBlock) {
        const uint64_t key = key_at(i);
        if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
      }
      __syncthreads();
      select_digit(hist, wt, shift, &s_prefix, &s_mask, &s_need, &s_done);
    }
  };
  if (in_lds) {
    for (int t = tid; t < M; t += kTopkBlock) {
      const uint64_t key = topk_key(scores, t);
      const uint32_t top = (uint32_t)(key >> 56);
      if (top > d1) {
        const uint32_t pos = atomicAdd(&s_count, 1u);
        if (pos < (uint32_t)NWPRT_MAX_K) sel[pos] = key;
      } else if (top == d1) {
        cand[atomicAdd(&s_ncand, 1u)] = key;
      }
    }
    __syncthreads();
    const int nc = (int)s_ncand;
    resolve(nc, [&](int i) { return cand[i]; });
    const uint64_t prefix = s_prefix, mask = s_mask;
    for (int i = tid; i < nc; i += kTopkBlock) {
      const uint64_t key = cand[i];
      if ((key & mask) >= prefix) {
        const uint32_t pos = atomicAdd(&s_count, 1u);
        if (pos < (uint32_t)NWPRT_MAX_K) sel[pos] = key;
      }
    }
  } else {
    resolve(M, [&](int t) { return topk_key(scores, t); });
    const uint64_t prefix = s_prefix, mask = s_mask;
    for (int t = tid; t < M; t += kTopkBlock) {
      const uint64_t key = topk_key(scores, t);
      if ((key & mask) >= prefix) {
        const uint32_t pos = atomicAdd(&s_count, 1u);
        if (pos < (uint32_t)NWPRT_MAX_K) sel[pos] = key;
      }
    }
  }
  __syncthreads();
  // Rank sort (descending) of the K kept keys: keys are unique, so key i goes
  // to position #{j : key_j > key_i}.  All lanes read the same sel[j] at a
  // time (an LDS broadcast); K <= 1024 = one key per thread.
  const int cnt = (int)s_count < K ? (int)s_count : K;
  for (int i = tid; i < cnt; i += kTopkBlock) {
    const uint64_t key = sel[i];
    int rank = 0;
    for (int j = 0; j < cnt; ++j) rank += sel[j] > key ? 1 : 0;
    sorted[rank] = key;
  }
  __syncthreads();
  for (int i = tid; i < K; i += kTopkBlock) {
    uint32_t idx = 0xFFFFFFFFu - (uint32_t)(sorted[i] & 0xFFFFFFFFull);
    if (idx >= (uint32_t)M) idx = 0;  // unreachable: keys are unique, exactly K are kept
    out_idx[i] = (int32_t)idx;
    for (int c = 0; c < 3; ++c) out_str[3 * i + c] = coef[3 * (size_t)idx + c];
  }
}

// -------------------------------------------------------- selection prep --

// One workgroup per selected wavelet k.  Thread t = (m-block b, lane):
// lane holds row m = 16 b + lane % 16 and input slots 8 g .. 8 g + 7
// (g = lane / 16) of the first layer's A operand, pre-multiplied by the
// wavelet feature row: W1[m][f] * wfeat[f][c_k].  This moves the per-pair
// Hadamard product of Renderer/tinycudann/utils_kernels.h:113-114 onto the
// weights, once per (relight, k).
__global__ __launch_bounds__(256) void prepare_selection_kernel(
    const float *__restrict__ w1, const float *__restrict__ wfeat, int n_coef,
    const int32_t *__restrict__ indices, int elem, uint16_t *__restrict__ mk,
    float *__restrict__ wsel, uint16_t *__restrict__ mk32, int planar,
    const float *__restrict__ proj) {
  __shared__ float wk[32], pj[32 * 32];
  const int k = blockIdx.x;
  const int t = threadIdx.x;
  const int c = indices[k];
  if (t < 32) wk[t] = wfeat[(size_t)t * n_coef + c];
  if (planar && proj)
    for (int i = t; i < 32 * 32; i += 256) pj[i] = proj[i];
  __syncthreads();
  // First-layer weight of output row m at contraction slot j: W1[m][j] w[j]
  // for row-layout features; for planar ones the raw feature planar_feature(j)
  // with the projection folded in, sum_i W1[m][i] w[i] P[i][f].
  auto weight = [&](int m, int j) -> float {
    if (!planar) return w1[m * 32 + j] * wk[j];
    const int f = planar_feature(j);
    if (!proj) return w1[m * 32 + f] * wk[f];
    float acc = 0.0f;
#pragma unroll 8
    for (int i = 0; i < 32; ++i) acc = fmaf(w1[m * 32 + i] * wk[i], pj[i * 32 + f], acc);
    return acc;
  };
  if (mk32) {
    // 32x32x16 A operand: block b = t / 128 (rows 32 b + lane % 32), K step
    // s = (t / 64) % 2, lane holds slots 16 s + 8 (lane / 32) + [0, 8).
    const int b = t >> 7, s = (t >> 6) & 1, lane = t & 63;
    const int m = 32 * b + (lane & 31);
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = weight(m, 16 * s + 8 * (lane >> 5) + j);
    dev::u32x4 out;
#pragma unroll
    for (int d = 0; d < 4; ++d) out[d] = dev::pack_bf16(v[2 * d], v[2 * d + 1]);
    *(dev::u32x4 *)(mk32 + (((size_t)k * 4 + b * 2 + s) * 64 + lane) * 8) = out;
  }
  if (mk) {
    const int b = t >> 6, lane = t & 63, g = lane >> 4;
    const int m = 16 * b + (lane & 15);
    float v[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) v[s] = weight(m, 8 * g + s);
    dev::u32x4 out;
#pragma unroll
    for (int d = 0; d < 4; ++d)
      out[d] = elem == (int)Elem::BF16 ? dev::pack_bf16(v[2 * d], v[2 * d + 1])
                                       : dev::pack_f16(v[2 * d], v[2 * d + 1]);
    *(dev::u32x4 *)(mk + (((size_t)k * 4 + b) * 64 + lane) * 8) = out;
  }
  if (wsel && t < 32) wsel[(size_t)k * 32 + t] = wk[t];
}

// Yaw rotation of a cubemap about +y: output texel (f, r, c) takes the
// source texel nearest to R^T d, where d is the texel centre's direction and
// R rotates by the yaw angle (x' = cos x + sin z, z' = -sin x + cos z).
// Face and (u, v) conventions of EnvUVToDirec / EnvDirecToUV
// (Renderer/device_functions/cubeMap.cu:23-115); row 0 is v = 0.  The
// direction arithmetic is FP32 with every operation rounded on its own
// (contraction is off in this file), so the texel choice is reproducible bit
// for bit on the host.
__device__ __forceinline__ void uv_to_dir(float u, float v, int f, float &x, float &y, float &z) {
  const float uc = 2.0f * u - 1.0f, vc = 2.0f * v - 1.0f;
  switch (f) {
    case 0: x = 1.0f; y = vc; z = uc; break;
    case 1: x = -1.0f; y = vc; z = -uc; break;
    case 2: x = uc; y = 1.0f; z = vc; break;
    case 3: x = uc; y = -1.0f; z = -vc; break;
    case 4: x = uc; y = vc; z = -1.0f; break;
    default: x = -uc; y = vc; z = 1.0f; break;
  }
}

__device__ __forceinline__ int dir_to_uv(float x, float y, float z, float &u, float &v) {
  const float ax = fabsf(x), ay = fabsf(y), az = fabsf(z);
  float m = 1.0f, uc = 0.0f, vc = 0.0f;
  int f = 0;
  // The reference tests the faces in this order and keeps the last match.
  if (x > 0 && ax >= ay && ax >= az) { m = ax; uc = z; vc = y; f = 0; }
  if (!(x > 0) && ax >= ay && ax >= az) { m = ax; uc = -z; vc = y; f = 1; }
  if (y > 0 && ay >= ax && ay >= az) { m = ay; uc = x; vc = z; f = 2; }
  if (!(y > 0) && ay >= ax && ay >= az) { m = ay; uc = x; vc = -z; f = 3; }
  if (!(z > 0) && az >= ax && az >= ay) { m = az; uc = x; vc = y; f = 4; }
  if (z > 0 && az >= ax && az >= ay) { m = az; uc = -x; vc = y; f = 5; }
  u = 0.5f * (uc / m + 1.0f);
  v = 0.5f * (vc / m + 1.0f);
  return f;
}

__global__ __launch_bounds__(256) void rotate_cubemap_kernel(const float *__restrict__ src,
                                                             float *__restrict__ dst, int n,
                                                             float cs, float sn) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 6 * n * n) return;
  const int f = t / (n * n), r = (t / n) % n, c = t % n;
  float x, y, z;
  uv_to_dir(((float)c + 0.5f) / (float)n, ((float)r + 0.5f) / (float)n, f, x, y, z);
  const float xs = cs * x - sn * z;
  const float zs = sn * x + cs * z;
  float u, v;
  const int fs = dir_to_uv(xs, y, zs, u, v);
  const int cs_ = min(n - 1, max(0, (int)floorf(u * (float)n)));
  const int rs = min(n - 1, max(0, (int)floorf(v * (float)n)));
  const float *p = src + (((size_t)fs * n + rs) * n + cs_) * 3;
  float *q = dst + (size_t)t * 3;
  q[0] = p[0];
  q[1] = p[1];
  q[2] = p[2];
}

}  // namespace

void launch_rotate_cubemap(const float *src, float *dst, int n, double yaw, hipStream_t s) {
  require(src && dst && n >= 1 && n <= NWPRT_MAX_FACE_SIZE, NWPRT_ERR_INVALID_ARGUMENT,
          "rotate: bad cubemap");
  require(src != dst, NWPRT_ERR_INVALID_ARGUMENT, "rotate: source and destination must differ");
  const int m = 6 * n * n;
  hipLaunchKernelGGL(rotate_cubemap_kernel, dim3((m + 255) / 256), dim3(256), 0, s, src, dst, n,
                     (float)std::cos(yaw), (float)std::sin(yaw));
  hip_check(hipGetLastError(), "rotate_cubemap_kernel launch");
}

size_t envmap_work_bytes(int n, int n_maps) {
  const size_t m = 6ull * n * n;
  const size_t topk = (m + (m + kScoreKeys - 1) / kScoreKeys * 256) * 4;
  const int T = n < kHaarTile ? n : kHaarTile;
  const size_t haar = (size_t)n_maps * 6 * (n / T) * (n / T) * 3 * 8;
  return topk > haar ? topk : haar;
}

void launch_haar(const float *faces, size_t map_stride, float *coef, int n, int n_maps,
                 void *work, hipStream_t s) {
  require(n >= 2 && n <= NWPRT_MAX_FACE_SIZE && (n & (n - 1)) == 0 && n_maps >= 1 && work &&
              n_maps <= 65535,
          NWPRT_ERR_INVALID_ARGUMENT, "haar: bad face size or map count");
  const int T = n < kHaarTile ? n : kHaarTile, tpf = n / T;
  hipLaunchKernelGGL(haar_tile_kernel, dim3(tpf * tpf, 6, n_maps), dim3(kHaarTileBlock), 0, s,
                     faces, map_stride, coef, n, (double *)work);
  hip_check(hipGetLastError(), "haar_tile_kernel launch");
  hipLaunchKernelGGL(haar_top_kernel, dim3(6, n_maps), dim3(64), 0, s, (const double *)work,
                     map_stride, coef, n);
  hip_check(hipGetLastError(), "haar_top_kernel launch");
}

void launch_topk(const float *coef, int n, int k, void *work, int32_t *indices,
                 float *strengths, hipStream_t s) {
  require(k >= 1 && k <= NWPRT_MAX_K && k <= 6 * n * n, NWPRT_ERR_INVALID_ARGUMENT,
          "topk: K must be in [1, min(1024, 6 N^2)]");
  require(n >= 2 && (n & (n - 1)) == 0 && work, NWPRT_ERR_INVALID_ARGUMENT,
          "topk: face size must be a power of two");
  const int m = 6 * n * n, n_wg = (m + kScoreKeys - 1) / kScoreKeys;
  uint32_t *scores = (uint32_t *)work, *wg_hist = scores + m;
  hipLaunchKernelGGL(score_kernel, dim3(n_wg), dim3(kScoreBlock), 0, s, coef, n, scores, wg_hist);
  hip_check(hipGetLastError(), "score_kernel launch");
  hipLaunchKernelGGL(topk_kernel, dim3(1), dim3(kTopkBlock), 0, s, coef, n, k, scores, wg_hist,
                     n_wg, indices, strengths);
  hip_check(hipGetLastError(), "topk_kernel launch");
}

void launch_prepare_selection(const float *w1, const float *wfeat, int n_coef,
                              const int32_t *indices, int k, Elem mk_elem, uint16_t *mk,
                              float *wsel, hipStream_t s, uint16_t *mk32, bool planar,
                              const float *projection) {
  require(k >= 1 && k <= NWPRT_MAX_K, NWPRT_ERR_INVALID_ARGUMENT, "prepare: bad K");
  require(mk == nullptr || mk_elem == Elem::BF16 || mk_elem == Elem::F16,
          NWPRT_ERR_INVALID_ARGUMENT, "prepare: MFMA tiles must be BF16 or FP16");
  hipLaunchKernelGGL(prepare_selection_kernel, dim3(k), dim3(256), 0, s, w1, wfeat, n_coef,
                     indices, (int)mk_elem, mk, wsel, mk32, (int)planar,
                     planar ? projection : nullptr);
  hip_check(hipGetLastError(), "prepare_selection_kernel launch");
}

}  // namespace nwprt
