/*
 * oracle.c -- FP64 CPU restatement of the relighting path (checker only;
 * see oracle.h for who may load it and how each function is pinned).
 *
 * Built with -ffp-contract=off: every FP64 expression below is evaluated
 * exactly as written, and the few FP32 steps that must match the GPU bit for
 * bit use explicit fmaf() where the kernel fuses.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------ hash grid -- */

/* Level table, in FP32 exactly as the product's host code documents it:
 * scale_l = exp2f(l * log2f(b)) * N_min - 1, res_l = ceilf(scale_l) + 1,
 * dense iff res_l^3 <= T, size_l = min(next_multiple(res_l^3, 8), T).
 * (tiny-cuda-nn's grid conventions behind Renderer/tinycudann/prt_network.h:45;
 * the library itself is absent from the reference tree.) */
void orc_hashgrid_levels(int n_levels, int log2_hashmap_size, int base_resolution,
                         float per_level_scale, float *scale, uint32_t *resolution,
                         uint32_t *dense, uint32_t *size, uint32_t *offset) {
  const uint64_t T = 1ull << log2_hashmap_size;
  const float log2_b = log2f(per_level_scale);
  uint64_t off = 0;
  for (int l = 0; l < n_levels; ++l) {
    const float s = exp2f((float)l * log2_b) * (float)base_resolution - 1.0f;
    const uint32_t r = (uint32_t)ceilf(s) + 1u;
    const uint64_t cube = (uint64_t)r * r * r;
    uint64_t n = cube <= T ? (cube + 7) / 8 * 8 : T;
    if (n > T) n = T;
    scale[l] = s;
    resolution[l] = r;
    dense[l] = cube <= T;
    size[l] = (uint32_t)n;
    offset[l] = (uint32_t)off;
    off += n;
  }
  offset[n_levels] = (uint32_t)off;
}

/* Multiresolution hash encoding of one position (Mueller et al. 2022,
 * "Instant Neural Graphics Primitives", Sec. 3): per level, the grid
 * position fmaf(x, scale, 0.5) in FP32, its floor and fraction, the 8
 * surrounding vertices, each mapped to an entry by the dense linear index
 * x + y res + z res^2 (dense levels) or by the spatial hash
 * x * 1 ^ y * 2654435761 ^ z * 805459861 (hashed levels), modulo the level
 * size; trilinear weights and the blend in FP64.  indices: [L][8] (may be
 * NULL), features: [L * F]. */
void orc_hash_encode(const orc_scene *sc, const float pos[3], uint32_t *indices,
                     double *features) {
  const int L = sc->n_levels, F = sc->n_features;
  for (int l = 0; l < L; ++l) {
    uint32_t g[3];
    float fr[3];
    for (int d = 0; d < 3; ++d) {
      const float p = fmaf(pos[d], sc->scale[l], 0.5f);
      const float fl = floorf(p);
      g[d] = (uint32_t)(int)fl;
      fr[d] = p - fl;
    }
    for (int j = 0; j < F; ++j) features[l * F + j] = 0.0;
    for (int c = 0; c < 8; ++c) {
      const uint32_t cx = g[0] + (c & 1), cy = g[1] + ((c >> 1) & 1), cz = g[2] + ((c >> 2) & 1);
      uint32_t idx;
      if (sc->dense[l]) {
        const uint32_t r = sc->resolution[l];
        idx = (cx + cy * r + cz * r * r) % sc->size[l];
      } else {
        idx = (cx ^ (cy * 2654435761u) ^ (cz * 805459861u)) % sc->size[l];
      }
      const double wx = (c & 1) ? (double)fr[0] : 1.0 - (double)fr[0];
      const double wy = (c & 2) ? (double)fr[1] : 1.0 - (double)fr[1];
      const double wz = (c & 4) ? (double)fr[2] : 1.0 - (double)fr[2];
      const double w = wx * wy * wz;
      const float *e = sc->table + ((size_t)sc->offset[l] + idx) * F;
      for (int j = 0; j < F; ++j) features[l * F + j] += w * (double)e[j];
      if (indices) indices[l * 8 + c] = idx;
    }
  }
}

/* ----------------------------------------------------------------- haar -- */

/* One 2x2 step of the orthonormal non-standard Haar transform; a = x[2i][2j],
 * b = x[2i][2j+1], c = x[2i+1][2j], d = x[2i+1][2j+1].  Row step then
 * column step with 1/sqrt(2) each == division by 2.  The operation order is
 * part of the specification (the GPU kernel uses the same). */
static void haar2x2(double a, double b, double c, double d, double *s, double *h, double *v,
                    double *dg) {
  *s = ((a + b) + (c + d)) * 0.5;
  *h = ((a - b) + (c - d)) * 0.5;
  *v = ((a + b) - (c + d)) * 0.5;
  *dg = ((a - b) - (c - d)) * 0.5;
}

/* Non-standard (Mallat pyramid) 2-D Haar transform of one N x N face
 * channel, elements `stride` floats apart, row 0 = bottom row (as faces are
 * stored after the reference's row flip, Renderer/HDRloader.h:21-25).  At a
 * level whose LL block is s x s (hs = s/2) the details of block (i, j) go to
 * (i, hs+j), (hs+i, j), (hs+i, hs+j); the last LL (face sum / N) goes to
 * (0, 0).  Scaling coefficients stay FP64; each output is rounded to FP32
 * once. */
void orc_haar_face(const float *src, int n, int stride, float *dst) {
  const int half = n / 2;
  double *cur = (double *)malloc(sizeof(double) * (size_t)half * half);
  double *nxt = (double *)malloc(sizeof(double) * (size_t)half * half);
  double s, h, v, dg;
  for (int i = 0; i < half; ++i)
    for (int j = 0; j < half; ++j) {
      haar2x2(src[((size_t)(2 * i) * n + 2 * j) * stride], src[((size_t)(2 * i) * n + 2 * j + 1) * stride],
              src[((size_t)(2 * i + 1) * n + 2 * j) * stride],
              src[((size_t)(2 * i + 1) * n + 2 * j + 1) * stride], &s, &h, &v, &dg);
      cur[i * half + j] = s;
      dst[((size_t)i * n + half + j) * stride] = (float)h;
      dst[((size_t)(half + i) * n + j) * stride] = (float)v;
      dst[((size_t)(half + i) * n + half + j) * stride] = (float)dg;
    }
  for (int sz = half; sz >= 2; sz /= 2) {
    const int hs = sz / 2;
    for (int i = 0; i < hs; ++i)
      for (int j = 0; j < hs; ++j) {
        haar2x2(cur[(2 * i) * half + 2 * j], cur[(2 * i) * half + 2 * j + 1],
                cur[(2 * i + 1) * half + 2 * j], cur[(2 * i + 1) * half + 2 * j + 1], &s, &h, &v,
                &dg);
        nxt[i * half + j] = s;
        dst[((size_t)i * n + hs + j) * stride] = (float)h;
        dst[((size_t)(hs + i) * n + j) * stride] = (float)v;
        dst[((size_t)(hs + i) * n + hs + j) * stride] = (float)dg;
      }
    double *t = cur;
    cur = nxt;
    nxt = t;
  }
  dst[0] = (float)cur[0];
  free(cur);
  free(nxt);
}

/* faces / coef: [6][N][N][3]. */
void orc_envmap_haar(const float *faces, int n, float *coef) {
  for (int f = 0; f < 6; ++f)
    for (int c = 0; c < 3; ++c)
      orc_haar_face(faces + (size_t)f * n * n * 3 + c, n, 3, coef + (size_t)f * n * n * 3 + c);
}

/* --------------------------------------------------------- yaw rotation -- */

/* Face and (u, v) conventions of EnvUVToDirec / EnvDirecToUV
 * (Renderer/device_functions/cubeMap.cu:99-115, 23-96), in FP32: this file
 * is compiled without contraction, so each operation rounds once, as the
 * product's explicitly rounded device arithmetic does. */
static void orc_uv_to_dir(float u, float v, int f, float d[3]) {
  const float uc = 2.0f * u - 1.0f, vc = 2.0f * v - 1.0f;
  switch (f) {
    case 0: d[0] = 1.0f; d[1] = vc; d[2] = uc; break;
    case 1: d[0] = -1.0f; d[1] = vc; d[2] = -uc; break;
    case 2: d[0] = uc; d[1] = 1.0f; d[2] = vc; break;
    case 3: d[0] = uc; d[1] = -1.0f; d[2] = -vc; break;
    case 4: d[0] = uc; d[1] = vc; d[2] = -1.0f; break;
    default: d[0] = -uc; d[1] = vc; d[2] = 1.0f; break;
  }
}

static int orc_dir_to_uv(const float d[3], float *u, float *v) {
  const float x = d[0], y = d[1], z = d[2];
  const float ax = fabsf(x), ay = fabsf(y), az = fabsf(z);
  float m = 1.0f, uc = 0.0f, vc = 0.0f;
  int f = 0;
  /* Tested in the reference's order; the last matching face wins. */
  if (x > 0 && ax >= ay && ax >= az) { m = ax; uc = z; vc = y; f = 0; }
  if (!(x > 0) && ax >= ay && ax >= az) { m = ax; uc = -z; vc = y; f = 1; }
  if (y > 0 && ay >= ax && ay >= az) { m = ay; uc = x; vc = z; f = 2; }
  if (!(y > 0) && ay >= ax && ay >= az) { m = ay; uc = x; vc = -z; f = 3; }
  if (!(z > 0) && az >= ax && az >= ay) { m = az; uc = x; vc = y; f = 4; }
  if (z > 0 && az >= ax && az >= ay) { m = az; uc = -x; vc = y; f = 5; }
  *u = 0.5f * (uc / m + 1.0f);
  *v = 0.5f * (vc / m + 1.0f);
  return f;
}

/* Yaw rotation about +y: texel (f, r, c) of dst takes the src texel nearest
 * to R^T d, d the direction of its centre ((c + 0.5) / N, (r + 0.5) / N), with
 * x' = cos x - sin z, z' = sin x + cos z (cos, sin of yaw in FP64, rounded to
 * FP32; the rest in FP32). */
void orc_rotate_cubemap(const float *src, int n, double yaw, float *dst) {
  const float cs = (float)cos(yaw), sn = (float)sin(yaw);
  for (int f = 0; f < 6; ++f)
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) {
        float d[3], e[3], u, v;
        orc_uv_to_dir(((float)c + 0.5f) / (float)n, ((float)r + 0.5f) / (float)n, f, d);
        e[0] = cs * d[0] - sn * d[2];
        e[1] = d[1];
        e[2] = sn * d[0] + cs * d[2];
        const int fs = orc_dir_to_uv(e, &u, &v);
        int cc = (int)floorf(u * (float)n), rr = (int)floorf(v * (float)n);
        cc = cc < 0 ? 0 : (cc > n - 1 ? n - 1 : cc);
        rr = rr < 0 ? 0 : (rr > n - 1 ? n - 1 : rr);
        memcpy(dst + (((size_t)f * n + r) * n + c) * 3, src + (((size_t)fs * n + rr) * n + cc) * 3,
               3 * sizeof(float));
      }
}

/* ---------------------------------------------------------------- top-K -- */

/* Area-weighted score of coefficient t: |luminance| with the weights and the
 * unfused FP32 evaluation of Renderer/EnvironmentMap.cpp:149-151, times the
 * relative support area of its basis function (4^-j for a detail at Mallat
 * position with max(row, col) in [2^j, 2^(j+1)), 1 for the DC term).
 * Subnormal, zero and NaN scores are 0. */
float orc_score(const float *coef, int t, int n) {
  const float r = coef[3 * (size_t)t], g = coef[3 * (size_t)t + 1], b = coef[3 * (size_t)t + 2];
  const float lum = (r * 0.212671f + g * 0.715160f) + b * 0.072169f;
  const int q = t % (n * n);
  const int row = q / n, col = q % n;
  const int m = row > col ? row : col;
  int j = 0;
  while ((2 << j) <= m) ++j;
  float s = ldexpf(fabsf(lum), -2 * j);
  if (!(s >= FLT_MIN)) s = 0.0f;
  return s;
}

typedef struct {
  float score;
  int32_t index;
} scored;

static int by_score_then_index(const void *pa, const void *pb) {
  const scored *a = (const scored *)pa, *b = (const scored *)pb;
  if (a->score != b->score) return a->score > b->score ? -1 : 1;
  return a->index < b->index ? -1 : (a->index > b->index);
}

/* The K largest scores, ties to the lower flat index, in that order;
 * strengths are the RGB coefficients at those indices (the layout read at
 * Renderer/EnvironmentMap.cpp:56-67). */
void orc_topk(const float *coef, int n, int k, int32_t *indices, float *strengths) {
  const int m = 6 * n * n;
  scored *all = (scored *)malloc(sizeof(scored) * (size_t)m);
  for (int t = 0; t < m; ++t) {
    all[t].score = orc_score(coef, t, n);
    all[t].index = t;
  }
  qsort(all, (size_t)m, sizeof(scored), by_score_then_index);
  for (int i = 0; i < k; ++i) {
    indices[i] = all[i].index;
    for (int c = 0; c < 3; ++c) strengths[3 * i + c] = coef[3 * (size_t)all[i].index + c];
  }
  free(all);
}

/* -------------------------------------------------------------- relight -- */

static double softplus_d(double x) { return x > 0 ? x + log1p(exp(-x)) : log1p(exp(x)); }

/* Per-pixel rows 3..15 of the reference's per-pair input
 * (Renderer/tinycudann/utils_kernels.h:31-46 with the offsets of
 * prt_globals.h:4-13), in the FP32 the reference computes them in: normal
 * and reflection direction as stored (already v/2 + 0.5), roughness as
 * log(25 r + 1) * (1 / log 26) (the multiply-add fused as a CUDA build
 * contracts it), 2 kd - 1, 2 ks - 1; then three padding inputs. */
void orc_aux_inputs(const orc_gbuffer *gb, int p, float pad, float out[16]) {
  const float *n = gb->normal + 4 * (size_t)p, *r = gb->reflect + 4 * (size_t)p;
  const float *kd = gb->kd + 4 * (size_t)p, *ks = gb->ks + 4 * (size_t)p;
  const float rgh_den = 1.0f / logf(26.0f);
  for (int c = 0; c < 3; ++c) {
    out[c] = n[c];
    out[3 + c] = r[c];
    out[7 + c] = fmaf(2.0f, kd[c], -1.0f);
    out[10 + c] = fmaf(2.0f, ks[c], -1.0f);
  }
  out[6] = logf(fmaf(25.0f, gb->roughness[p], 1.0f)) * rgh_den;
  out[13] = out[14] = out[15] = pad;
}

/* Relight of selected pixels in FP64.  Per pixel p with env_flag set
 * (Renderer/tinycudann/utils_kernels.h:72):
 *   h = hash(pos) (orc_hash_encode), h' = proj h (or h)
 *   for each selected wavelet k with flat index c_k
 *     z = h' (.) wfeat[:, c_k]              (utils_kernels.h:113-114)
 *     T = out_act(W_out relu(... relu(W1 z)))
 *     rgb += max(light[k] * T, 0) per channel (utils_kernels.h:74-78)
 * indirect = rgb, with_direct = rgb + direct (utils_kernels.h:80-88).
 * With aux inputs, waux * aux(p) joins the second layer's pre-activation
 * (the folded cmp_net input [out_prd | cmp_enc], prt_network.h:113-139).
 * Background pixels give 0.  transport (optional): [n][K][3]. */
void orc_relight_pixels(const orc_scene *sc, const int32_t *sel, const float *light, int k,
                        const orc_gbuffer *gb, const int32_t *pixels, int n_pixels,
                        double *indirect, double *with_direct, double *transport, int n_threads) {
  const int Fd = sc->feature_dim, H = sc->hidden, NH = sc->n_hidden_layers;
#pragma omp parallel for schedule(dynamic, 4) num_threads(n_threads > 0 ? n_threads : 1)
  for (int q = 0; q < n_pixels; ++q) {
    const int p = pixels ? pixels[q] : q;
    double rgb[3] = {0, 0, 0};
    const int hit = gb->env_flag == NULL || gb->env_flag[p] != 0;
    if (hit) {
      double feat[64], hp[64], z[64], a[128], b[128], b2[128];
      for (int o = 0; o < H; ++o) b2[o] = 0;
      if (sc->aux) {
        float ax[16];
        orc_aux_inputs(gb, p, sc->aux_pad, ax);
        for (int o = 0; o < H; ++o) {
          double acc = 0;
          for (int i = 0; i < 16; ++i) acc += (double)sc->waux[o * 16 + i] * (double)ax[i];
          b2[o] = acc;
        }
      }
      orc_hash_encode(sc, gb->position + 4 * (size_t)p, NULL, feat);
      for (int o = 0; o < Fd; ++o) {
        if (sc->proj) {
          double acc = 0;
          for (int i = 0; i < Fd; ++i) acc += (double)sc->proj[o * Fd + i] * feat[i];
          hp[o] = acc;
        } else {
          hp[o] = feat[o];
        }
      }
      for (int kk = 0; kk < k; ++kk) {
        const int c = sel[kk];
        for (int f = 0; f < Fd; ++f) z[f] = hp[f] * (double)sc->wfeat[(size_t)f * sc->n_coef + c];
        const float *W = sc->layers;
        for (int o = 0; o < H; ++o) {
          double acc = 0;
          for (int i = 0; i < Fd; ++i) acc += (double)W[o * Fd + i] * z[i];
          a[o] = acc > 0 ? acc : 0;
        }
        W += (size_t)H * Fd;
        for (int l = 1; l < NH; ++l) {
          for (int o = 0; o < H; ++o) {
            double acc = l == 1 ? b2[o] : 0;
            for (int i = 0; i < H; ++i) acc += (double)W[o * H + i] * a[i];
            b[o] = acc > 0 ? acc : 0;
          }
          memcpy(a, b, sizeof(double) * (size_t)H);
          W += (size_t)H * H;
        }
        for (int ch = 0; ch < 3; ++ch) {
          double acc = 0;
          for (int i = 0; i < H; ++i) acc += (double)W[ch * H + i] * a[i];
          const double t = sc->softplus ? softplus_d(acc) : acc;
          const double term = (double)light[3 * kk + ch] * t;
          rgb[ch] += term > 0 ? term : 0;
          if (transport) transport[((size_t)q * k + kk) * 3 + ch] = t;
        }
      }
    } else if (transport) {
      for (int kk = 0; kk < 3 * k; ++kk) transport[(size_t)q * k * 3 + kk] = 0;
    }
    for (int ch = 0; ch < 3; ++ch) {
      indirect[3 * q + ch] = rgb[ch];
      with_direct[3 * q + ch] = rgb[ch] + (double)gb->direct[4 * (size_t)p + ch];
    }
  }
}

/* --------------------------------------------------------------- tonemap -- */

/* Renderer/toneMap.cu:58-70: y = clamp01(1.055 * (x + 1e-6)^(1/2.4) - 0.055),
 * byte = (uint)(y * 255.9), packed R | G << 8 | B << 16 | 255 << 24.  The
 * byte is the step function of the curve taken in FP64: T_b (b = 1..255) is
 * the smallest FP32 x with 1.055 * pow((double)x + 1e-6, 1/2.4) - 0.055 >=
 * b / 255.9, and byte(x) = #{b : x >= T_b} (NaN and x < T_1 give 0).  The
 * reference's FP32 powf agrees with it except within a few ulps of a T_b. */
static double tm_curve(float x) { return 1.055 * pow((double)x + 1e-6, 1.0 / 2.4) - 0.055; }

static float tm_table[256];
static int tm_ready = 0;

static void tm_build(void) {
  if (tm_ready) return;
  for (int b = 1; b < 256; ++b) {
    const double want = (double)b / 255.9;
    /* start from the inverse curve, then walk to the first float that reaches b */
    float x = (float)(pow((want + 0.055) / 1.055, 2.4) - 1e-6);
    while (tm_curve(x) < want) x = nextafterf(x, INFINITY);
    while (tm_curve(nextafterf(x, -INFINITY)) >= want) x = nextafterf(x, -INFINITY);
    tm_table[b] = x;
  }
  tm_ready = 1;
}

void orc_tonemap_thresholds(float *out255) {
  tm_build();
  for (int b = 1; b < 256; ++b) out255[b - 1] = tm_table[b];
}

static uint32_t tm_byte(float x) {
  uint32_t n = 0;
  for (int b = 1; b < 256; ++b) n += (x >= tm_table[b]);
  return n;
}

uint32_t orc_tonemap(float r, float g, float b) {
  tm_build();
  return tm_byte(r) | (tm_byte(g) << 8) | (tm_byte(b) << 16) | (255u << 24);
}

/* Guided edge-avoiding a-trous denoiser (Dammertz, Sewtz, Hanika, Lensch,
 * HPG 2010), the filter that replaces the reference's closed learned
 * denoiser behind OptixBackend::denoise (Renderer/OptixBackend.cpp:841-906):
 * colour, normal and albedo are float4 [h][w]; kappa = 2^-mean(log2 Y) over
 * pixels with Y = 0.2126 R + 0.7152 G + 0.0722 B > 1e-8 (1 if none);
 * pass i (step 2^i) replaces each pixel's RGB by the weighted mean of the
 * in-image 5 x 5 taps with weight h(dx) h(dy) exp(-(kappa^2 |dc|^2 4^i / sc^2
 * + |dn|^2 / sn^2 + |da|^2 / sa^2)), h = (1, 4, 6, 4, 1) / 16; the output
 * is blend * colour + (1 - blend) * the last pass, alpha = colour alpha.
 * FP64 throughout.  out: double [h][w][4]. */
void orc_denoise(const float *color, const float *normal, const float *albedo, int w, int h,
                 int iterations, double sigma_color, double sigma_normal, double sigma_albedo,
                 double blend, double *out) {
  const size_t P = (size_t)w * h;
  double s = 0.0, cnt = 0.0;
  for (size_t p = 0; p < P; ++p) {
    const double y = 0.2126 * (double)color[4 * p] + 0.7152 * (double)color[4 * p + 1] +
                     0.0722 * (double)color[4 * p + 2];
    if (y > 1e-8) {
      s += log2(y);
      cnt += 1.0;
    }
  }
  const double kappa = cnt > 0.0 ? exp2(-s / cnt) : 1.0;
  double *cur = (double *)malloc(P * 3 * sizeof(double));
  double *nxt = (double *)malloc(P * 3 * sizeof(double));
  for (size_t p = 0; p < P; ++p)
    for (int c = 0; c < 3; ++c) cur[3 * p + c] = color[4 * p + c];
  static const double kh[5] = {1.0 / 16, 4.0 / 16, 6.0 / 16, 4.0 / 16, 1.0 / 16};
  for (int i = 0; i < iterations; ++i) {
    const int step = 1 << i;
    const double ic = kappa * kappa * (double)(1 << (2 * i)) / (sigma_color * sigma_color);
    const double in_ = 1.0 / (sigma_normal * sigma_normal);
    const double ia = 1.0 / (sigma_albedo * sigma_albedo);
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) {
        const size_t p = (size_t)y * w + x;
        double acc[3] = {0.0, 0.0, 0.0}, ws = 0.0;
        for (int dy = -2; dy <= 2; ++dy) {
          const int qy = y + dy * step;
          if (qy < 0 || qy >= h) continue;
          for (int dx = -2; dx <= 2; ++dx) {
            const int qx = x + dx * step;
            if (qx < 0 || qx >= w) continue;
            const size_t q = (size_t)qy * w + qx;
            double dc = 0.0, dn = 0.0, da = 0.0;
            for (int c = 0; c < 3; ++c) {
              const double a = cur[3 * q + c] - cur[3 * p + c];
              const double b = (double)normal[4 * q + c] - (double)normal[4 * p + c];
              const double e = (double)albedo[4 * q + c] - (double)albedo[4 * p + c];
              dc += a * a;
              dn += b * b;
              da += e * e;
            }
            const double wt = kh[dx + 2] * kh[dy + 2] * exp(-(ic * dc + in_ * dn + ia * da));
            for (int c = 0; c < 3; ++c) acc[c] += wt * cur[3 * q + c];
            ws += wt;
          }
        }
        for (int c = 0; c < 3; ++c) nxt[3 * p + c] = acc[c] / ws;
      }
    double *t = cur;
    cur = nxt;
    nxt = t;
  }
  for (size_t p = 0; p < P; ++p) {
    for (int c = 0; c < 3; ++c)
      out[4 * p + c] = blend * (double)color[4 * p + c] + (1.0 - blend) * cur[3 * p + c];
    out[4 * p + 3] = color[4 * p + 3];
  }
  free(cur);
  free(nxt);
}

/* ------------------------------------------------------------------------ */
/* G-buffer producer: first hits of primary rays, brute force over every
 * triangle in FP64.  Follows __raygen__fovCamera (Renderer/device_functions/
 * camera.cu:82-131, 233-270) with the jitter at its mean, RenderEngine::
 * setCamera (Renderer/RenderEngine.cpp:246-267), __closesthit__radiance
 * (Renderer/device_functions/devicePrograms.cu:47-109, 156-164, 295-308),
 * __miss__radiance (devicePrograms.cu:517-564) and EnvDirecToUV
 * (Renderer/device_functions/cubeMap.cu:21-92) sampled bilinearly with wrap
 * addressing (Renderer/textures.cpp:31-36). */

static void g_cross(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static double g_dot(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static void g_norm(double *a) {
  const double l = sqrt(g_dot(a, a));
  a[0] /= l;
  a[1] /= l;
  a[2] /= l;
}

static void g_env(const float *faces, int n, const double *d, double *rgb) {
  const double ax = fabs(d[0]), ay = fabs(d[1]), az = fabs(d[2]);
  double m = ax, uc = d[2], vc = d[1];
  int f = 0;
  if (d[0] > 0 && ax >= ay && ax >= az) { m = ax; uc = d[2]; vc = d[1]; f = 0; }
  if (!(d[0] > 0) && ax >= ay && ax >= az) { m = ax; uc = -d[2]; vc = d[1]; f = 1; }
  if (d[1] > 0 && ay >= ax && ay >= az) { m = ay; uc = d[0]; vc = d[2]; f = 2; }
  if (!(d[1] > 0) && ay >= ax && ay >= az) { m = ay; uc = d[0]; vc = -d[2]; f = 3; }
  if (!(d[2] > 0) && az >= ax && az >= ay) { m = az; uc = d[0]; vc = d[1]; f = 4; }
  if (d[2] > 0 && az >= ax && az >= ay) { m = az; uc = -d[0]; vc = d[1]; f = 5; }
  if (!(m > 0)) { /* no direction: no radiance, as the product's lookup */
    rgb[0] = rgb[1] = rgb[2] = 0.0;
    return;
  }
  const double u = 0.5 * (uc / m + 1.0), v = 0.5 * (vc / m + 1.0);
  const double x = u * n - 0.5, y = v * n - 0.5;
  const double fx = floor(x), fy = floor(y), wx = x - fx, wy = y - fy;
  int x0 = ((int)fx % n + n) % n, y0 = ((int)fy % n + n) % n;
  const int x1 = (x0 + 1) % n, y1 = (y0 + 1) % n;
  const float *face = faces + (size_t)f * n * n * 3;
  for (int c = 0; c < 3; ++c) {
    const double p00 = face[((size_t)y0 * n + x0) * 3 + c], p01 = face[((size_t)y0 * n + x1) * 3 + c];
    const double p10 = face[((size_t)y1 * n + x0) * 3 + c], p11 = face[((size_t)y1 * n + x1) * 3 + c];
    const double top = p00 + wx * (p01 - p00), bot = p10 + wx * (p11 - p10);
    rgb[c] = top + wy * (bot - top);
  }
}

/* Closest hit of o + t d over all triangles, t > 0.01, ties to the lower
 * index.  Returns the triangle or -1; *margin is lowered to the smallest
 * robustness figure of this ray: the barycentric distance to an edge of any
 * triangle that could take the hit, the relative gap to the next hit, and
 * |det| / (|e1| |e2|) of the hit triangle (grazing incidence). */
static int g_closest(const orc_mesh *m, const double *o, const double *d, double *t_out,
                     double *u_out, double *v_out, double *margin) {
  const double tmin = 1e-2;
  double best = 1e300, second = 1e300, bu = 0, bv = 0;
  int hit = -1;
  for (int64_t t = 0; t < m->n_tris; ++t) {
    const float *A = m->vertices + 3 * (size_t)m->indices[3 * t];
    const float *B = m->vertices + 3 * (size_t)m->indices[3 * t + 1];
    const float *C = m->vertices + 3 * (size_t)m->indices[3 * t + 2];
    double e1[3], e2[3], s[3], p[3], q[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = (double)B[k] - A[k];
      e2[k] = (double)C[k] - A[k];
      s[k] = o[k] - A[k];
    }
    g_cross(d, e2, p);
    const double det = g_dot(e1, p);
    if (det == 0.0) continue;
    const double u = g_dot(s, p) / det;
    g_cross(s, e1, q);
    const double v = g_dot(d, q) / det;
    const double tt = g_dot(e2, q) / det;
    if (!(tt > tmin)) continue;
    const double w = 1.0 - u - v;
    const double e = fmin(u, fmin(v, w));
    if (e >= 0.0) {
      if (tt < best) {
        second = best;
        best = tt;
        hit = (int)t;
        bu = u;
        bv = v;
      } else if (tt < second) {
        second = tt;
      }
    }
  }
  /* Second pass for the margin: triangles whose edge passes near the ray at
   * a distance that could take (or lose) the hit. */
  for (int64_t t = 0; t < m->n_tris; ++t) {
    const float *A = m->vertices + 3 * (size_t)m->indices[3 * t];
    const float *B = m->vertices + 3 * (size_t)m->indices[3 * t + 1];
    const float *C = m->vertices + 3 * (size_t)m->indices[3 * t + 2];
    double e1[3], e2[3], s[3], p[3], q[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = (double)B[k] - A[k];
      e2[k] = (double)C[k] - A[k];
      s[k] = o[k] - A[k];
    }
    g_cross(d, e2, p);
    const double det = g_dot(e1, p);
    if (det == 0.0) continue;
    const double u = g_dot(s, p) / det;
    g_cross(s, e1, q);
    const double v = g_dot(d, q) / det;
    const double tt = g_dot(e2, q) / det;
    if (!(tt > tmin * (1.0 - 1e-6))) continue;
    if (hit >= 0 && tt > best * (1.0 + 1e-6) + 1e-9) continue;
    const double e = fabs(fmin(u, fmin(v, 1.0 - u - v)));
    if (e < *margin) *margin = e;
    if (fabs(tt - tmin) / tmin < *margin) *margin = fabs(tt - tmin) / tmin;
  }
  if (hit >= 0) {
    if (second < 1e300 && (second - best) / best < *margin) *margin = (second - best) / best;
    const float *A = m->vertices + 3 * (size_t)m->indices[3 * hit];
    const float *B = m->vertices + 3 * (size_t)m->indices[3 * hit + 1];
    const float *C = m->vertices + 3 * (size_t)m->indices[3 * hit + 2];
    double e1[3], e2[3], p[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = (double)B[k] - A[k];
      e2[k] = (double)C[k] - A[k];
    }
    g_cross(d, e2, p);
    const double g = fabs(g_dot(e1, p)) / (sqrt(g_dot(e1, e1)) * sqrt(g_dot(e2, e2)));
    if (g < *margin) *margin = g;
  }
  *t_out = best;
  *u_out = bu;
  *v_out = bv;
  return hit;
}

static void g_direct(const orc_mesh *m, const double *tab, int n, const float *light,
                     const double *p, const double *d, const double *ns, const double *ng,
                     const float *mat, int samples, uint32_t seed, uint32_t pixel, double *out);

static void g_tint(const double *c, double *t);
static const float *g_material_row(const orc_mesh *m, int mat, const int *tri, double u,
                                   double v, float *buf);

static void g_cast(const orc_mesh *m, const float *camera, int w, int h, int max_bounce,
                   const float *bounds, const float *env_faces, int env_n, double *position,
                   double *normal, double *reflect, double *kd, double *ks, double *roughness,
                   uint8_t *env_flag, int32_t *prim, double *margin, const float *light,
                   const double *tab, int light_n, int samples, uint32_t seed, double *direct) {
  double dir[3], up[3], hz[3], vt[3];
  for (int a = 0; a < 3; ++a) {
    dir[a] = (double)camera[3 + a] - camera[a];
    up[a] = camera[6 + a];
  }
  g_norm(dir);
  g_cross(dir, up, hz);
  g_norm(hz);
  for (int a = 0; a < 3; ++a) hz[a] *= (double)w / (double)h;
  g_cross(hz, dir, vt);
  g_norm(vt);
  const double th = tan(0.5 * 0.8575560548920328);
  const double cx = w / 2.0, cy = h / 2.0, fx = 0.5 * w / th, fy = 0.5 * h / th;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t px = 0; px < (int64_t)w * h; ++px) {
    const int ix = (int)(px % w), iy = (int)(px / w);
    const double sx = (ix - cx) / fx, sy = (iy - cy) / fy;
    double d[3], o[3];
    for (int a = 0; a < 3; ++a) {
      d[a] = dir[a] + sx * hz[a] + sy * vt[a];
      o[a] = camera[a];
    }
    g_norm(d);
    double pos[3] = {0, 0, 0}, nrm[3] = {0, 0, 0}, rfl[3] = {0, 0, 0};
    double kdv[3] = {0, 0, 0}, ksv[3] = {0, 0, 0}, rough = 0.0, mg = 1.0;
    double dir_l[3] = {0, 0, 0};
    int flag = 0, first = -1;
    for (int depth = 1;; ++depth) {
      double t, u, v;
      const int hit = g_closest(m, o, d, &t, &u, &v, &mg);
      if (hit < 0) {
        for (int a = 0; a < 3; ++a) {
          nrm[a] = -d[a];
          rfl[a] = -d[a];
          pos[a] = -100000.0;
        }
        if (depth == 1 && env_faces) g_env(env_faces, env_n, d, kdv);
        for (int a = 0; a < 3; ++a) dir_l[a] = kdv[a];
        rough = 1.0;
        break;
      }
      const int mat = m->material_ids ? m->material_ids[hit] : 0;
      const int *tri = m->indices + 3 * (size_t)hit;
      float mbuf[ORC_MAT_FLOATS];
      const float *mr = g_material_row(m, mat, tri, u, v, mbuf);
      double ns[3], ng[3];
      {
        const float *A = m->vertices + 3 * (size_t)tri[0];
        const float *B = m->vertices + 3 * (size_t)tri[1];
        const float *C = m->vertices + 3 * (size_t)tri[2];
        double e1[3], e2[3];
        for (int a = 0; a < 3; ++a) {
          e1[a] = (double)B[a] - A[a];
          e2[a] = (double)C[a] - A[a];
        }
        g_cross(e1, e2, ng);
        g_norm(ng);
      }
      if (m->normals) {
        const double ww = 1.0 - u - v;
        for (int a = 0; a < 3; ++a)
          ns[a] = ww * m->normals[3 * (size_t)tri[0] + a] + u * m->normals[3 * (size_t)tri[1] + a] +
                  v * m->normals[3 * (size_t)tri[2] + a];
      } else {
        const float *A = m->vertices + 3 * (size_t)tri[0];
        const float *B = m->vertices + 3 * (size_t)tri[1];
        const float *C = m->vertices + 3 * (size_t)tri[2];
        double e1[3], e2[3];
        for (int a = 0; a < 3; ++a) {
          e1[a] = (double)B[a] - A[a];
          e2[a] = (double)C[a] - A[a];
        }
        g_cross(e1, e2, ns);
      }
      g_norm(ns);
      double p[3], r[3];
      for (int a = 0; a < 3; ++a) p[a] = o[a] + d[a] * t + 0.001;
      const double nd = g_dot(ns, d);
      for (int a = 0; a < 3; ++a) r[a] = d[a] - 2.0 * nd * ns[a];
      g_norm(r);
      if ((int)mr[7] == 3) { /* mirror: the path continues */
        if (depth >= max_bounce) break;
        for (int a = 0; a < 3; ++a) {
          o[a] = p[a];
          d[a] = r[a];
        }
        continue;
      }
      for (int a = 0; a < 3; ++a) {
        pos[a] = p[a];
        nrm[a] = ns[a];
        rfl[a] = r[a];
        kdv[a] = mr[a];
        ksv[a] = mr[3 + a];
      }
      if ((int)mr[7] == 1) { /* DisneyBSDF's PACKINFO (disneyBSDF.cu:10-25) */
        double ct[3];
        g_tint(kdv, ct);
        const double eta = mr[18];
        const double f0 = (eta - 1.0) * (eta - 1.0) / ((eta + 1.0) * (eta + 1.0));
        for (int a = 0; a < 3; ++a)
          ksv[a] = mr[11] * f0 * (1.0 - mr[9]) * ((1.0 - mr[12]) + mr[12] * ct[a]) + mr[9] * kdv[a];
      }
      rough = mr[6];
      flag = 1;
      first = hit;
      if (direct && light)
        g_direct(m, tab, light_n, light, p, d, ns, ng, mr, samples, seed, (uint32_t)px, dir_l);
      break;
    }
    for (int a = 0; a < 3; ++a) {
      position[4 * px + a] = (pos[a] - bounds[a]) / ((double)bounds[3 + a] - bounds[a]);
      normal[4 * px + a] = nrm[a] * 0.5 + 0.5;
      reflect[4 * px + a] = rfl[a] * 0.5 + 0.5;
      kd[4 * px + a] = kdv[a];
      ks[4 * px + a] = ksv[a];
    }
    position[4 * px + 3] = normal[4 * px + 3] = reflect[4 * px + 3] = 1.0;
    kd[4 * px + 3] = ks[4 * px + 3] = 1.0;
    roughness[px] = rough;
    env_flag[px] = (uint8_t)flag;
    prim[px] = first;
    margin[px] = mg;
    if (direct)
      for (int a = 0; a < 4; ++a) direct[4 * px + a] = a < 3 ? dir_l[a] : 1.0;
  }
}

void orc_gbuffer_cast(const orc_mesh *m, const float *camera, int w, int h, int max_bounce,
                      const float *bounds, const float *env_faces, int env_n, double *position,
                      double *normal, double *reflect, double *kd, double *ks, double *roughness,
                      uint8_t *env_flag, int32_t *prim, double *margin) {
  g_cast(m, camera, w, h, max_bounce, bounds, env_faces, env_n, position, normal, reflect, kd, ks,
         roughness, env_flag, prim, margin, NULL, NULL, 0, 0, 0, NULL);
}

/* ---- direct light --------------------------------------------------------
 * The closest-hit program's direct term in wavelet mode (maximumBounce 1,
 * Renderer/device_functions/devicePrograms.cu:165-280) for the cookTorrance
 * (Renderer/device_functions/cookTorrance.cu:26-195), roughPlastic and
 * simple_ggx BSDFs (type from the material row), replaying the
 * product's counter-based random stream: per sample, an environment sample
 * (Renderer/device_functions/cubeMap.cu:115-186 over the tables of
 * Renderer/EnvironmentMap.cpp:128-190 and Renderer/table_dist.cpp:4-130)
 * and a BSDF sample, each shadow-tested by brute force, power-heuristic
 * weights.  Directions, BSDF and weights in FP64; the tables' texel weights
 * in the reference's FP32 operations; random numbers are the same FP32
 * values the product draws. */

float orc_rng_uniform(uint32_t seed, uint32_t pixel, uint32_t counter) {
  uint64_t z = (((uint64_t)seed << 32) | pixel) + 0x9E3779B97F4A7C15ull * ((uint64_t)counter + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (float)(uint32_t)(z >> 40) * 5.9604644775390625e-8f;
}

static float g_lumf(float x, float y, float z) {
  const float a = x * 0.212671f, b = y * 0.715160f, c = z * 0.072169f;
  return (a + b) + c;
}

static size_t g_face_doubles(int n) {
  return (size_t)n * (n + 1) + (size_t)n * n + (size_t)(n + 1) + (size_t)n;
}

size_t orc_env_table_doubles(int n) { return 19 + 6 * g_face_doubles(n); }

static double g_area(double x, double y) { return atan2(x * y, sqrt(x * x + y * y + 1.0)); }

void orc_env_tables(const float *faces, int n, double *tab) {
  const float k = (2.0f * 3.14159265358979323846f) / 3.0f;
  for (int f = 0; f < 6; ++f) {
    double *base = tab + 19 + (size_t)f * g_face_doubles(n);
    double *cdf_rows = base, *pdf_rows = base + (size_t)n * (n + 1);
    double *cdf_m = pdf_rows + (size_t)n * n, *pdf_m = cdf_m + (n + 1);
    double *row_int = (double *)malloc(sizeof(double) * n), lum = 0.0;
    for (int r = 0; r < n; ++r) {
      double *cr = cdf_rows + (size_t)r * (n + 1), *pr = pdf_rows + (size_t)r * n, ls = 0.0;
      cr[0] = 0.0;
      for (int c = 0; c < n; ++c) {
        /* TexelCoordSolidAngle((r + 0.5) / n, (c + 0.5) / n, n) */
        const double U = 2.0 * ((r + 0.5) / n + 0.5) / n - 1.0;
        const double V = 2.0 * ((c + 0.5) / n + 0.5) / n - 1.0;
        const double iv = 1.0 / n;
        const double sa = g_area(U - iv, V - iv) - g_area(U - iv, V + iv) -
                          g_area(U + iv, V - iv) + g_area(U + iv, V + iv);
        const float saf = (float)sa;
        const float *t = faces + (((size_t)f * n + r) * n + c) * 3;
        const float qx = (fabsf(t[0]) * saf) / k, qy = (fabsf(t[1]) * saf) / k,
                    qz = (fabsf(t[2]) * saf) / k;
        const double fv = (double)g_lumf(qx, qy, qz);
        pr[c] = fv;
        cr[c + 1] = cr[c] + fv;
        ls += (double)g_lumf(t[0], t[1], t[2]);
      }
      const double integral = cr[n];
      for (int c = 0; c < n; ++c) {
        if (integral > 0) {
          cr[c] /= integral;
          pr[c] /= integral;
        } else {
          cr[c] = pr[c] = 0.0;
        }
      }
      if (!(integral > 0)) cr[n] = 0.0;
      row_int[r] = cr[n];
      lum += ls;
    }
    cdf_m[0] = 0.0;
    for (int y = 0; y < n; ++y) cdf_m[y + 1] = cdf_m[y] + row_int[y];
    const double total = cdf_m[n];
    for (int y = 0; y < n; ++y) {
      if (total > 0) {
        cdf_m[y] /= total;
        pdf_m[y] = row_int[y] / total;
      } else {
        pdf_m[y] = 1.0 / n;
        cdf_m[y] = (double)y / n;
      }
    }
    cdf_m[n] = 1.0;
    for (int r = 0; r < n; ++r) {
      double *cr = cdf_rows + (size_t)r * (n + 1);
      cr[n] = cr[n - 1] > 0 ? 1.0 : 0.0;
    }
    tab[13 + f] = lum;
    free(row_int);
  }
  double total = 0.0, fr[6], cdf[7];
  for (int i = 0; i < 6; ++i) total += tab[13 + i];
  for (int i = 0; i < 6; ++i) fr[i] = total > 0 ? tab[13 + i] / total : 0.0;
  cdf[0] = 0.0;
  for (int i = 0; i < 6; ++i) cdf[i + 1] = cdf[i] + fr[i];
  const double t = cdf[6];
  for (int i = 0; i < 6; ++i) {
    tab[i] = t > 0 ? fr[i] / t : 1.0 / 6.0;
    tab[6 + i] = t > 0 ? cdf[i] / t : i / 6.0;
  }
  tab[12] = t > 0 ? t : 1.0;
}

static int g_first_above(const double *cdf, int count, float r) {
  for (int i = 0; i < count; ++i)
    if (cdf[i] > (double)r) return i;
  return count;
}

static int g_clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static void g_face_to_dir(int f, double u, double v, double *d) {
  const double uc = 2.0 * u - 1.0, vc = 2.0 * v - 1.0;
  switch (f) {
    case 0: d[0] = 1; d[1] = vc; d[2] = uc; break;
    case 1: d[0] = -1; d[1] = vc; d[2] = -uc; break;
    case 2: d[0] = uc; d[1] = 1; d[2] = vc; break;
    case 3: d[0] = uc; d[1] = -1; d[2] = -vc; break;
    case 4: d[0] = uc; d[1] = vc; d[2] = -1; break;
    default: d[0] = -uc; d[1] = vc; d[2] = 1; break;
  }
  g_norm(d);
}

static void g_env_sample(const double *tab, int n, float r0, float r1, float r2, double *d) {
  const int f = g_clampi(g_first_above(tab + 6, 6, r0) - 1, 0, 5);
  const double *base = tab + 19 + (size_t)f * g_face_doubles(n);
  const double *cdf_m = base + (size_t)n * (n + 1) + (size_t)n * n;
  const int y = g_clampi(g_first_above(cdf_m, n, r2) - 1, 0, n - 1);
  double dy = (double)r2 - cdf_m[y];
  if (cdf_m[y + 1] - cdf_m[y] > 0) dy /= (cdf_m[y + 1] - cdf_m[y]);
  const double *cdf = base + (size_t)y * (n + 1);
  const int x = g_clampi(g_first_above(cdf, n, r1) - 1, 0, n - 1);
  double dx = (double)r1 - cdf[x];
  if (cdf[x + 1] - cdf[x] > 0) dx /= (cdf[x + 1] - cdf[x]);
  /* the sampled (u, v) is returned as FP32 (cubeMap.cu:164) */
  g_face_to_dir(f, (double)(float)((x + dx) / n), (double)(float)((y + dy) / n), d);
}

static int g_dir_to_face(const double *d, double *u, double *v) {
  const double ax = fabs(d[0]), ay = fabs(d[1]), az = fabs(d[2]);
  double m = ax, uc = d[2], vc = d[1];
  int f = 0;
  if (d[0] > 0 && ax >= ay && ax >= az) { m = ax; uc = d[2]; vc = d[1]; f = 0; }
  if (!(d[0] > 0) && ax >= ay && ax >= az) { m = ax; uc = -d[2]; vc = d[1]; f = 1; }
  if (d[1] > 0 && ay >= ax && ay >= az) { m = ay; uc = d[0]; vc = d[2]; f = 2; }
  if (!(d[1] > 0) && ay >= ax && ay >= az) { m = ay; uc = d[0]; vc = -d[2]; f = 3; }
  if (!(d[2] > 0) && az >= ax && az >= ay) { m = az; uc = d[0]; vc = d[1]; f = 4; }
  if (d[2] > 0 && az >= ax && az >= ay) { m = az; uc = -d[0]; vc = d[1]; f = 5; }
  *u = 0.5 * (uc / m + 1.0);
  *v = 0.5 * (vc / m + 1.0);
  return f;
}

static double g_env_pdf(const double *tab, int n, const double *d) {
  double u, v;
  const int f = g_dir_to_face(d, &u, &v);
  const double *base = tab + 19 + (size_t)f * g_face_doubles(n);
  const double *pdf_rows = base + (size_t)n * (n + 1);
  const double *pdf_m = pdf_rows + (size_t)n * n + (n + 1);
  const int x = (int)fmin(fmax(u * n, 0.0), (double)(n - 1));
  const int y = (int)fmin(fmax(v * n, 0.0), (double)(n - 1));
  return tab[f] * pdf_m[y] * pdf_rows[(size_t)y * n + x] * n * n;
}

static double g_lum(const double *c) {
  return c[0] * 0.212671 + c[1] * 0.715160 + c[2] * 0.072169;
}

/* Orthonormal frame of Frisvad (2012) around unit n. */
static void g_frame(const double *n, double *t, double *b) {
  if (n[2] < -1.0 + 1e-6) {
    t[0] = 0; t[1] = -1; t[2] = 0;
    b[0] = -1; b[1] = 0; b[2] = 0;
  } else {
    const double a = 1.0 / (1.0 + n[2]), c = -n[0] * n[1] * a;
    t[0] = 1.0 - n[0] * n[0] * a; t[1] = c; t[2] = -n[0];
    b[0] = c; b[1] = 1.0 - n[1] * n[1] * a; b[2] = -n[1];
  }
}
static void g_to_local(const double *d, const double *n0, double *out) {
  double n[3] = {n0[0], n0[1], n0[2]}, t[3], b[3];
  g_norm(n);
  g_frame(n, t, b);
  out[0] = g_dot(d, t);
  out[1] = g_dot(d, b);
  out[2] = g_dot(d, n);
}
static void g_to_world(const double *d, const double *n0, double *out) {
  double n[3] = {n0[0], n0[1], n0[2]}, t[3], b[3];
  g_norm(n);
  g_frame(n, t, b);
  if (!(n[2] < -1.0 + 1e-6)) {
    g_norm(t);
    g_norm(b);
  }
  for (int a = 0; a < 3; ++a) out[a] = t[a] * d[0] + b[a] * d[1] + n[a] * d[2];
}
static double g_gtr2(const double *h, double ax, double ay) {
  const double x = h[0] / ax, y = h[1] / ay, l2 = x * x + y * y + h[2] * h[2];
  return 1.0 / (M_PI * ax * ay * l2 * l2);
}
static double g_smith(const double *v, double ax, double ay) {
  const double lam =
      (-1.0 + sqrt(1.0 + (v[0] * v[0] * ax * ax + v[1] * v[1] * ay * ay) / (v[2] * v[2]))) / 2.0;
  return 1.0 / (1.0 + lam);
}

typedef struct {
  double kd[3], ks[3], rough, ns[3], ng[3];
  int type; /* 0 roughPlastic, 1 DisneyBSDF, 2 cookTorrance, 4 simple_ggx */
  /* DisneyBSDF parameters (Renderer/Material.h:45-55); eta also for
   * roughPlastic */
  double st, metallic, subsurface, specular, spec_tint, aniso, sheen, sheen_tint, clearcoat,
      cc_gloss, eta;
} g_surface;

/* Material row (ORC_MAT_FLOATS values): kd[3], ks[3], roughness, type, then
 * specular_transmission, metallic, subsurface, specular, specular_tint,
 * anisotropic, sheen, sheen_tint, clearcoat, clearcoat_gloss, eta. */
/* Texture sampling for the G-buffer producer (Renderer/textures.cpp:8-47,
 * EightBitTexture::toCUDA): the IEC 61966-2-1 sRGB decode of a byte, and
 * tex2D with normalized coordinates, wrap addressing and bilinear
 * filtering, all in FP64. */
static double g_srgb(int b) {
  const double c = b / 255.0;
  return c <= 0.04045 ? c / 12.92 : pow((c + 0.055) / 1.055, 2.4);
}

void orc_tex_sample(const orc_mesh *m, int t, double u, double v, double *out) {
  const int w = m->tex_w[t], h = m->tex_h[t];
  const uint8_t *px = m->tex_texels + m->tex_offset[t];
  double fu = u - floor(u), fv = v - floor(v);
  if (!(fu >= 0.0 && fu < 1.0)) fu = 0.0;
  if (!(fv >= 0.0 && fv < 1.0)) fv = 0.0;
  const double x = fu * w - 0.5, y = fv * h - 0.5, x0f = floor(x), y0f = floor(y);
  const double ax = x - x0f, ay = y - y0f;
  int x0 = (int)x0f, y0 = (int)y0f;
  if (x0 < 0) x0 += w;
  if (y0 < 0) y0 += h;
  const int x1 = (x0 + 1) % w, y1 = (y0 + 1) % h;
  const int xs[4] = {x0, x1, x0, x1}, ys[4] = {y0, y0, y1, y1};
  const double ws[4] = {(1 - ax) * (1 - ay), ax * (1 - ay), (1 - ax) * ay, ax * ay};
  for (int c = 0; c < 4; ++c) {
    double acc = 0.0;
    for (int k = 0; k < 4; ++k) {
      const int b = px[((size_t)ys[k] * w + xs[k]) * 4 + c];
      acc += ws[k] * (c < 3 ? g_srgb(b) : b / 255.0);
    }
    out[c] = acc;
  }
}

/* The material row of a hit with its textured fields read at the hit's
 * texture coordinates (1 - u - v) t0 + u t1 + v t2 (devicePrograms.cu:
 * 66-71, 121-147); the row itself when the mesh has no textures. */
static const float *g_material_row(const orc_mesh *m, int mat, const int *tri, double u,
                                   double v, float *buf) {
  const float *mr = m->materials + ORC_MAT_FLOATS * (size_t)mat;
  if (!m->texcoords || !m->mat_tex) return mr;
  memcpy(buf, mr, ORC_MAT_FLOATS * sizeof(float));
  const double w = 1.0 - u - v;
  double tc[2];
  for (int a = 0; a < 2; ++a)
    tc[a] = w * m->texcoords[2 * (size_t)tri[0] + a] + u * m->texcoords[2 * (size_t)tri[1] + a] +
            v * m->texcoords[2 * (size_t)tri[2] + a];
  for (int f = 0; f < ORC_TEX_FIELDS; ++f) {
    const int t = m->mat_tex[(size_t)mat * ORC_TEX_FIELDS + f];
    if (t < 0) continue;
    double x[4];
    orc_tex_sample(m, t, tc[0], tc[1], x);
    if (f == 0 || f == 1)
      for (int a = 0; a < 3; ++a) buf[3 * f + a] = (float)x[a];
    else if (f == 2)
      buf[6] = (float)x[0];
    else
      buf[8 + (f - 3)] = (float)x[0];
  }
  return buf;
}

static void g_surface_of(const float *mat, const double *ns, const double *ng, g_surface *s) {
  for (int a = 0; a < 3; ++a) {
    s->kd[a] = mat[a];
    s->ks[a] = mat[3 + a];
    s->ns[a] = ns[a];
    s->ng[a] = ng[a];
  }
  s->rough = mat[6];
  s->type = (int)mat[7];
  s->st = mat[8];
  s->metallic = mat[9];
  s->subsurface = mat[10];
  s->specular = mat[11];
  s->spec_tint = mat[12];
  s->aniso = mat[13];
  s->sheen = mat[14];
  s->sheen_tint = mat[15];
  s->clearcoat = mat[16];
  s->cc_gloss = mat[17];
  s->eta = mat[18];
}

static void g_shading_normal(const g_surface *s, const double *v, double *n) {
  const double sg = g_dot(s->ns, v) < 0 ? -1.0 : 1.0;
  for (int a = 0; a < 3; ++a) n[a] = sg * s->ns[a];
}

static void g_ct_eval(const g_surface *s, const double *l, const double *v, double *out) {
  out[0] = out[1] = out[2] = 0.0;
  if (g_dot(s->ng, v) < 0 || g_dot(s->ng, l) < 0) return;
  double n[3], h[3], hl[3], vl[3], ll[3];
  g_shading_normal(s, v, n);
  if (g_lum(s->ks) + g_lum(s->kd) <= 0) return;
  for (int a = 0; a < 3; ++a) h[a] = v[a] + l[a];
  g_norm(h);
  const double n_in = g_dot(n, v), n_out = g_dot(n, l);
  const double alpha = fmax(0.0025, s->rough * s->rough);
  if (g_dot(h, n) < 0)
    for (int a = 0; a < 3; ++a) h[a] = -h[a];
  const double h_out = g_dot(h, l);
  const double sch = pow(fmax(1.0 - h_out, 0.0), 5.0);
  g_to_local(h, n, hl);
  g_to_local(v, n, vl);
  g_to_local(l, n, ll);
  const double k = g_gtr2(hl, alpha, alpha) * g_smith(vl, alpha, alpha) *
                   g_smith(ll, alpha, alpha) / (4.0 * n_in * n_out);
  const double base = 1.05 * (1.0 - pow(1.0 - n_out, 5.0)) * (1.0 - pow(1.0 - n_in, 5.0)) / M_PI;
  for (int c = 0; c < 3; ++c) {
    const double f = s->ks[c] + (1.0 - s->ks[c]) * sch;
    out[c] = ((1.0 - s->ks[c]) * base * s->kd[c] + f * k) * n_out;
  }
}

/* Specular-lobe probability: by luminance (cookTorrance), or
 * clamp(1 - roughness, 0.05, 0.75) for simple_ggx
 * (Renderer/device_functions/simple_ggx.cu:54-55).  0: black BSDF. */
static int g_spec_probability(const g_surface *s, double *spec) {
  if (s->type == 4) {
    *spec = fmin(fmax(1.0 - (double)fminf(fmaxf((float)s->rough, 0.05f), 1.0f), 0.05), 0.75);
    *spec = (double)(float)*spec;
    return 1;
  }
  const double ls = g_lum(s->ks), lr = g_lum(s->kd);
  if (ls + lr <= 0) return 0;
  *spec = ls / (ls + lr);
  return 1;
}

static double g_ct_pdf(const g_surface *s, const double *l, const double *v) {
  if (g_dot(s->ng, v) < 0 || g_dot(s->ng, l) < 0) return 0.0;
  double n[3], h[3], hl[3], vl[3];
  g_shading_normal(s, v, n);
  for (int a = 0; a < 3; ++a) h[a] = v[a] + l[a];
  g_norm(h);
  const double n_h = g_dot(n, h), n_in = g_dot(n, v), n_out = g_dot(n, l);
  if (n_out <= 0 || n_h <= 0) return 0.0;
  double spec;
  if (!g_spec_probability(s, &spec)) return 0.0;
  const double r = fmin(fmax(s->rough, 0.05), 1.0), a = r * r;
  const double diff = 1.0 - spec;
  g_to_local(v, n, vl);
  g_to_local(h, n, hl);
  spec *= g_smith(vl, a, a) * g_gtr2(hl, a, a) / (4.0 * n_in);
  return spec + diff * n_out / M_PI;
}

/* Returns 0 for a black BSDF (the reference's (-2, -2, -2)). */
static int g_ct_sample(const g_surface *s, const double *v, float rw, float ru, float rv,
                       double *out) {
  double n[3], spec;
  g_shading_normal(s, v, n);
  if (!g_spec_probability(s, &spec)) return 0;
  if ((double)rw > spec) {
    const double phi = 2.0 * M_PI * ru, tmp = sqrt(fmin(fmax(1.0 - rv, 0.0), 1.0));
    const double loc[3] = {cos(phi) * tmp, sin(phi) * tmp, sqrt(fmin(fmax((double)rv, 0.0), 1.0))};
    g_to_world(loc, n, out);
    return 1;
  }
  double li[3];
  g_to_local(v, n, li);
  const int below = li[2] < 0;
  if (below)
    for (int a = 0; a < 3; ++a) li[a] = -li[a];
  const double r = fmin(fmax(s->rough, 0.05), 1.0), a = r * r;
  double hemi[3] = {a * li[0], a * li[1], li[2]};
  g_norm(hemi);
  const double rr = sqrt((double)ru), phi = 2.0 * M_PI * rv;
  const double t1 = rr * cos(phi);
  double t2 = rr * sin(phi);
  const double sm = (1.0 + hemi[2]) / 2.0;
  t2 = (1.0 - sm) * sqrt(1.0 - t1 * t1) + sm * t2;
  const double disk[3] = {t1, t2, sqrt(fmax(0.0, 1.0 - t1 * t1 - t2 * t2))};
  double hn[3], micro[3], hw[3];
  g_to_world(disk, hemi, hn);
  micro[0] = a * hn[0];
  micro[1] = a * hn[1];
  micro[2] = fmax(0.0, hn[2]);
  g_norm(micro);
  g_to_world(micro, n, hw);
  g_norm(hw);
  const double vh = g_dot(v, hw);
  for (int k = 0; k < 3; ++k) out[k] = -v[k] + 2.0 * vh * hw[k];
  g_norm(out);
  if (below)
    for (int k = 0; k < 3; ++k) out[k] = -out[k];
  return 1;
}

/* roughPlastic (Renderer/device_functions/roughPlastic.cu:25-174) at the
 * material's index of refraction eta (Renderer/Material.h:55, 94-97). */

static double g_fresnel(double n_dot_i, double eta) {
  const double t2 = 1.0 - (1.0 - n_dot_i * n_dot_i) / (eta * eta);
  if (t2 < 0) return 1.0;
  const double nt = sqrt(t2), ni = fabs(n_dot_i);
  const double rs = (ni - eta * nt) / (ni + eta * nt), rp = (eta * ni - nt) / (eta * ni + nt);
  return (rs * rs + rp * rp) / 2.0;
}

static void g_rp_eval(const g_surface *s, const double *l, const double *v, double *out) {
  out[0] = out[1] = out[2] = 0.0;
  double n[3], h[3], hl[3], vl[3], ll[3];
  g_shading_normal(s, v, n);
  for (int a = 0; a < 3; ++a) h[a] = v[a] + l[a];
  g_norm(h);
  const double n_h = g_dot(n, h), n_in = g_dot(n, v), n_out = g_dot(n, l);
  if (n_out <= 0 || n_h <= 0) return;
  const double r = fmin(fmax(s->rough, 0.01), 1.0), a = r * r;
  const double fo = g_fresnel(g_dot(h, l), s->eta), fi = g_fresnel(g_dot(h, v), s->eta);
  g_to_local(h, n, hl);
  g_to_local(v, n, vl);
  g_to_local(l, n, ll);
  const double k = g_smith(vl, a, a) * g_smith(ll, a, a) * fo * g_gtr2(hl, a, a) /
                   (4.0 * n_in * n_out);
  const double base = (1.0 - fo) * (1.0 - fi) / M_PI;
  for (int c = 0; c < 3; ++c) out[c] = (s->kd[c] * base + s->ks[c] * k) * n_out;
}

static double g_rp_pdf(const g_surface *s, const double *l, const double *v) {
  if (g_dot(s->ng, v) < 0 || g_dot(s->ng, l) < 0) return 0.0;
  double n[3], h[3], hl[3], vl[3];
  g_shading_normal(s, v, n);
  for (int a = 0; a < 3; ++a) h[a] = v[a] + l[a];
  g_norm(h);
  const double n_h = g_dot(n, h), n_in = g_dot(n, v), n_out = g_dot(n, l);
  if (n_out <= 0 || n_h <= 0) return 0.0;
  const double ls = g_lum(s->ks), lr = g_lum(s->kd);
  if (ls + lr <= 0) return 0.0;
  const double r = fmin(fmax(s->rough, 0.01), 1.0), a = r * r;
  double spec = ls / (ls + lr);
  const double diff = 1.0 - spec;
  g_to_local(v, n, vl);
  g_to_local(h, n, hl);
  spec *= g_smith(vl, a, a) * g_gtr2(hl, a, a) / (4.0 * n_in);
  return spec + diff * n_out / M_PI;
}

/* z1 chooses the lobe and is reused for the direction, as the reference
 * does.  Returns 0 for the reference's (-2, -2, -2). */
static int g_rp_sample(const g_surface *s, const double *v, float z1, float z2, double *out) {
  if (g_dot(s->ng, v) < 0) return 0;
  double n[3];
  g_shading_normal(s, v, n);
  const double ls = g_lum(s->ks), lr = g_lum(s->kd);
  if (ls + lr <= 0) return 0;
  const double spec = ls / (ls + lr);
  if ((double)z1 > spec) {
    const double phi = 2.0 * M_PI * z1, tmp = sqrt(fmin(fmax(1.0 - z2, 0.0), 1.0));
    const double loc[3] = {cos(phi) * tmp, sin(phi) * tmp, sqrt(fmin(fmax((double)z2, 0.0), 1.0))};
    g_to_world(loc, n, out);
    return 1;
  }
  double li[3];
  g_to_local(v, n, li);
  const double r = fmin(fmax(s->rough, 0.01), 1.0), a = r * r;
  double hemi[3] = {a * li[0], a * li[1], li[2]};
  g_norm(hemi);
  const double rr = sqrt((double)z1), phi = 2.0 * M_PI * z2;
  const double t1 = rr * cos(phi);
  double t2 = rr * sin(phi);
  const double sm = (1.0 + hemi[2]) / 2.0;
  t2 = (1.0 - sm) * sqrt(1.0 - t1 * t1) + sm * t2;
  const double disk[3] = {t1, t2, sqrt(fmax(0.0, 1.0 - t1 * t1 - t2 * t2))};
  double hn[3], micro[3], hw[3];
  g_to_world(disk, hemi, hn);
  micro[0] = a * hn[0];
  micro[1] = a * hn[1];
  micro[2] = fmax(0.0, hn[2]);
  g_norm(micro);
  g_to_world(micro, n, hw);
  g_norm(hw);
  const double vh = g_dot(v, hw);
  for (int k = 0; k < 3; ++k) out[k] = -v[k] + 2.0 * vh * hw[k];
  g_norm(out);
  return 1;
}

/* simple_ggx's value (Renderer/device_functions/simple_ggx.cu:74-126). */
static void g_sg_eval(const g_surface *s, const double *l, const double *v, double *out) {
  out[0] = out[1] = out[2] = 0.0;
  if (g_dot(s->ng, v) < 0 || g_dot(s->ng, l) < 0) return;
  double n[3], h[3];
  g_shading_normal(s, v, n);
  for (int a = 0; a < 3; ++a) h[a] = v[a] + l[a];
  g_norm(h);
  const double n_l = fmin(fmax(g_dot(n, l), 1e-6), 1.0), n_v = fmin(fmax(g_dot(n, v), 1e-6), 1.0);
  const double n_h = fmin(fmax(g_dot(n, h), 1e-6), 1.0), v_h = fmin(fmax(g_dot(h, v), 1e-6), 1.0);
  const double a = s->rough * s->rough, a2 = a * a, k = (a + 2.0 * s->rough + 1.0) / 8.0;
  const double fmi = (-5.55473 * v_h - 6.98316) * v_h;
  const double frac = (0.04 + 0.96 * pow(2.0, fmi)) * a2;
  const double nom0 = n_h * n_h * (a2 - 1.0) + 1.0;
  const double nom = fmin(fmax(4.0 * M_PI * nom0 * nom0 * (n_v * (1 - k) + k) * (n_l * (1 - k) + k),
                               1e-6), 4.0 * M_PI);
  for (int c = 0; c < 3; ++c) out[c] = frac / nom * n_l + s->kd[c] / M_PI * n_l;
}

/* DisneyBSDF (Renderer/device_functions/disneyBSDF.cu:44-492) with the
 * helpers of shaderUtils.h:55-213: value, pdf and sampler as the reference
 * writes them, including the diffuse pdf's cosine of the view direction and
 * the lobe probabilities that the refraction case leaves unscaled. */
typedef struct {
  double n[3]; /* shading normal flipped to the view's geometric side */
  double ax, ay, ac, cc, dw, mw, gw, cw;
} g_dlobes;

static void g_disney_lobes(const g_surface *s, const double *v, g_dlobes *d) {
  const double sg = g_dot(s->ns, v) * g_dot(s->ng, v) < 0 ? -1.0 : 1.0;
  for (int a = 0; a < 3; ++a) d->n[a] = sg * s->ns[a];
  d->cc = 0.25 * s->clearcoat;
  const double aspect = sqrt(1.0 - s->aniso * 0.9);
  d->ax = fmax(0.0001, s->rough * s->rough / aspect);
  d->ay = fmax(0.0001, s->rough * s->rough * aspect);
  d->ac = (1.0 - s->cc_gloss) * 0.1 + s->cc_gloss * 0.001;
  d->dw = (1.0 - s->metallic) * (1.0 - s->st);
  d->mw = 1.0 - s->st * (1.0 - s->metallic);
  d->gw = (1.0 - s->metallic) * s->st;
  d->cw = d->cc;
}

static double g_gtr1(double n_dot_h, double a) {
  const double a2 = a * a, t = 1.0 + (a2 - 1.0) * n_dot_h * n_dot_h;
  return (a2 - 1.0) / (M_PI * log(a2) * t);
}

static void g_tint(const double *c, double *t) {
  const double l = g_lum(c);
  for (int a = 0; a < 3; ++a) t[a] = l > 0 ? c[a] / l : 1.0;
}

/* G(v) G(l) D(h) in the frame of n with alphas (ax, ay). */
static double g_dg(const double *h, const double *v, const double *l, const double *n, double ax,
                   double ay, int with_l) {
  double hl[3], vl[3], ll[3];
  g_to_local(h, n, hl);
  g_to_local(v, n, vl);
  double r = g_gtr2(hl, ax, ay) * g_smith(vl, ax, ay);
  if (with_l) {
    g_to_local(l, n, ll);
    r *= g_smith(ll, ax, ay);
  }
  return r;
}

static void g_ds_eval(const g_surface *s, const double *l, const double *v, double *out) {
  g_dlobes d;
  g_disney_lobes(s, v, &d);
  const double *n = d.n;
  out[0] = out[1] = out[2] = 0.0;
  const int reflect = g_dot(s->ng, l) * g_dot(s->ng, v) > 0;
  if (reflect) {
    const double n_in = g_dot(n, v), n_out = g_dot(n, l);
    double h[3];
    for (int a = 0; a < 3; ++a) h[a] = v[a] + l[a];
    g_norm(h);
    if (g_dot(h, n) < 0)
      for (int a = 0; a < 3; ++a) h[a] = -h[a];
    const double n_h = g_dot(s->ns, h), h_in = g_dot(h, v), h_out = g_dot(h, l);
    if (g_dot(s->ng, v) >= 0 && g_dot(s->ng, l) >= 0 && n_out > 0) {
      const double fd90 = 0.5 + 2.0 * s->rough * h_out * h_out;
      const double so = pow(1.0 - n_out, 5.0), si = pow(1.0 - n_in, 5.0);
      const double sh = pow(1.0 - h_out, 5.0);
      const double base = (1.0 + (fd90 - 1.0) * so) * (1.0 + (fd90 - 1.0) * si);
      const double fss90 = h_out * h_out * s->rough;
      const double fss = (1.0 + (fss90 - 1.0) * so) * (1.0 + (fss90 - 1.0) * si);
      const double ss = 1.25 * (fss * (1.0 / (n_out + n_in) - 0.5) + 0.5);
      const double kdiff = (1.0 - s->st) * (1.0 - s->metallic) *
                           ((base * (1.0 - s->subsurface) + ss * s->subsurface) / M_PI) * n_out;
      double ct[3];
      g_tint(s->kd, ct);
      for (int a = 0; a < 3; ++a) {
        out[a] += s->kd[a] * kdiff;
        out[a] += ((1.0 - s->sheen_tint) + s->sheen_tint * ct[a]) * (1.0 - s->metallic) *
                  s->sheen * sh * n_out;
      }
      if (n_in > 0 && h_out > 0 && n_h > 0) {
        const double f0 = (s->eta - 1.0) * (s->eta - 1.0) / ((s->eta + 1.0) * (s->eta + 1.0));
        const double sw = 1.0 - s->st * (1.0 - s->metallic);
        const double fw = pow(fmax(1.0 - h_out, 0.0), 5.0);
        const double k = sw * g_dg(h, v, l, n, d.ax, d.ay, 1) / (4.0 * n_in);
        for (int a = 0; a < 3; ++a) {
          const double sc = (1.0 - s->spec_tint) + s->spec_tint * ct[a];
          const double c0 = s->specular * f0 * (1.0 - s->metallic) * sc + s->metallic * s->kd[a];
          out[a] += (c0 + (1.0 - c0) * fw) * k;
        }
      }
      if (n_in > 0 && n_h > 0) {
        const double fc = 0.04 + 0.96 * pow(fmax(1.0 - h_out, 0.0), 5.0);
        double vl[3], ll[3];
        g_to_local(v, n, vl);
        g_to_local(l, n, ll);
        const double gc = g_smith(vl, 0.25, 0.25) * g_smith(ll, 0.25, 0.25);
        const double c = d.cc * fc * g_gtr1(n_h, d.ac) * gc / (4.0 * n_in);
        for (int a = 0; a < 3; ++a) out[a] += c;
      }
    }
    const double eta = g_dot(s->ng, v) > 0 ? s->eta : 1.0 / s->eta;
    const double g = (1.0 - s->metallic) * s->st * g_fresnel(h_in, eta) *
                     g_dg(h, v, l, n, d.ax, d.ay, 1) / (4.0 * fabs(n_in));
    for (int a = 0; a < 3; ++a) out[a] += s->kd[a] * g;
    return;
  }
  const double eta = g_dot(s->ng, v) > 0 ? s->eta : 1.0 / s->eta;
  double h[3];
  for (int a = 0; a < 3; ++a) h[a] = v[a] + l[a] * eta;
  g_norm(h);
  if (g_dot(h, n) < 0)
    for (int a = 0; a < 3; ++a) h[a] = -h[a];
  const double h_in = g_dot(h, v), h_out = g_dot(h, l), sd = h_in + eta * h_out;
  const double t = (1.0 - s->metallic) * s->st *
                   ((1.0 / (eta * eta)) * (1.0 - g_fresnel(h_in, eta)) *
                    g_dg(h, v, l, n, d.ax, d.ay, 1) * eta * eta * fabs(h_out * h_in)) /
                   (fabs(g_dot(n, v)) * sd * sd);
  for (int a = 0; a < 3; ++a) out[a] = sqrt(s->kd[a]) * t;
}

static double g_ds_pdf(const g_surface *s, const double *l, const double *v) {
  g_dlobes d;
  g_disney_lobes(s, v, &d);
  const double *n = d.n;
  const int reflect = g_dot(s->ng, l) * g_dot(s->ng, v) > 0;
  const double tot = d.dw + d.mw + d.gw + d.cw;
  double dp = d.dw / tot, mp = d.mw / tot, gp = d.gw / tot, cp = d.cw / tot;
  if (g_dot(s->ng, v) < 0) {
    dp = mp = cp = 0.0;
    if (gp > 0) gp = 1.0;
  }
  const double eta = g_dot(s->ng, v) > 0 ? s->eta : 1.0 / s->eta;
  double h[3];
  if (reflect) {
    for (int a = 0; a < 3; ++a) h[a] = v[a] + l[a];
    g_norm(h);
    if (g_dot(h, n) < 0)
      for (int a = 0; a < 3; ++a) h[a] = -h[a];
    const double n_in = g_dot(n, v), n_h = g_dot(h, s->ns), h_in = g_dot(h, v),
                 h_out = g_dot(h, l);
    if (g_dot(s->ng, l) >= 0 && g_dot(s->ng, v) >= 0) {
      dp *= fmax(g_dot(n, v), 0.0) / M_PI;
      mp = n_in > 0 ? mp * g_dg(h, v, l, n, d.ax, d.ay, 0) / (4.0 * n_in) : 0.0;
      cp = (n_h > 0 && h_out > 0) ? cp * g_gtr1(n_h, d.ac) * n_h / (4.0 * h_out) : 0.0;
    }
    gp *= g_fresnel(h_in, eta) * g_dg(h, v, l, n, d.ax, d.ay, 0) / (4.0 * fabs(n_in));
  } else {
    for (int a = 0; a < 3; ++a) h[a] = v[a] + l[a] * eta;
    g_norm(h);
    if (g_dot(h, n) < 0)
      for (int a = 0; a < 3; ++a) h[a] = -h[a];
    const double h_in = g_dot(h, v), h_out = g_dot(h, l), sd = h_in + eta * h_out;
    const double dh = eta * eta * h_out / (sd * sd);
    gp *= (1.0 - g_fresnel(h_in, eta)) * g_dg(h, v, l, n, d.ax, d.ay, 0) *
          fabs(dh * h_in / g_dot(n, v));
  }
  return dp + mp + gp + cp;
}

/* Heitz's visible-normal sample with alphas (ax, ay); an incoming direction
 * below the surface is flipped (shaderUtils.h:160-213). */
static void g_vndf(const double *li0, double ax, double ay, float u1, float u2, double *m) {
  double li[3] = {li0[0], li0[1], li0[2]};
  const int below = li[2] < 0;
  if (below)
    for (int a = 0; a < 3; ++a) li[a] = -li[a];
  double hemi[3] = {ax * li[0], ay * li[1], li[2]}, disk[3], hn[3];
  g_norm(hemi);
  const double r = sqrt((double)u1), phi = 2.0 * M_PI * u2;
  const double t1 = r * cos(phi);
  double t2 = r * sin(phi);
  const double sm = (1.0 + hemi[2]) / 2.0;
  t2 = (1.0 - sm) * sqrt(1.0 - t1 * t1) + sm * t2;
  disk[0] = t1;
  disk[1] = t2;
  disk[2] = sqrt(fmax(0.0, 1.0 - t1 * t1 - t2 * t2));
  g_to_world(disk, hemi, hn);
  m[0] = ax * hn[0];
  m[1] = ay * hn[1];
  m[2] = fmax(0.0, hn[2]);
  g_norm(m);
  if (below)
    for (int a = 0; a < 3; ++a) m[a] = -m[a];
}

static void g_reflect_about(const double *v, const double *h, double *out) {
  const double vh = g_dot(v, h);
  for (int a = 0; a < 3; ++a) out[a] = -v[a] + 2.0 * vh * h[a];
  g_norm(out);
}

/* Glass lobe; total internal reflection returns `tir` (0: no sample, 1: the
 * zero vector). */
static int g_ds_glass(const g_surface *s, const g_dlobes *d, const double *v, double u, float ru,
                      float rv, int tir, double *out) {
  double li[3], m[3], h[3];
  g_to_local(v, s->ns, li);
  g_vndf(li, d->ax, d->ay, ru, rv, m);
  g_to_world(m, s->ns, h);
  if (g_dot(h, d->n) < 0)
    for (int a = 0; a < 3; ++a) h[a] = -h[a];
  const double h_in = g_dot(h, v);
  const double eta = g_dot(s->ng, v) > 0 ? s->eta : 1.0 / s->eta;
  if (u <= g_fresnel(h_in, eta)) {
    g_reflect_about(v, h, out);
    return 1;
  }
  const double hos = 1.0 - (1.0 - h_in * h_in) / (eta * eta);
  if (hos <= 0) {
    out[0] = out[1] = out[2] = 0.0;
    return tir;
  }
  if (h_in < 0)
    for (int a = 0; a < 3; ++a) h[a] = -h[a];
  const double ho = sqrt(hos);
  for (int a = 0; a < 3; ++a) out[a] = -v[a] / eta + (fabs(h_in) / eta - ho) * h[a];
  return 1;
}

static int g_ds_sample(const g_surface *s, const double *v, float rw, float ru, float rv,
                       double *out) {
  g_dlobes d;
  g_disney_lobes(s, v, &d);
  if (g_dot(s->ng, v) >= 0) {
    const double tot = d.dw + d.mw + d.gw + d.cw;
    const double dp = d.dw / tot, mp = d.mw / tot, gp = d.gw / tot;
    if (rw <= dp) {
      const double phi = 2.0 * M_PI * ru;
      const double tmp = sqrt(fmin(fmax(1.0 - rv, 0.0), 1.0));
      const double loc[3] = {cos(phi) * tmp, sin(phi) * tmp, sqrt(fmin(fmax((double)rv, 0.0), 1.0))};
      g_to_world(loc, d.n, out);
      return 1;
    }
    if (rw <= dp + mp) {
      double li[3], m[3], h[3];
      g_to_local(v, s->ns, li);
      g_vndf(li, d.ax, d.ay, ru, rv, m);
      g_to_world(m, s->ns, h);
      g_reflect_about(v, h, out);
      return 1;
    }
    if (rw <= dp + mp + gp) {
      if (gp <= 0) return 0;
      return g_ds_glass(s, &d, v, (rw - (dp + mp)) / gp, ru, rv, 0, out);
    }
    const double a2 = d.ac * d.ac;
    const double ch = sqrt(fmax(0.0, (1.0 - pow(a2, 1.0 - ru)) / (1.0 - a2)));
    const double sh = sqrt(fmax(1.0 - ch * ch, 0.0)), az = 2.0 * M_PI * rv;
    const double m[3] = {sh * cos(az), sh * sin(az), ch};
    double h[3];
    g_to_world(m, d.n, h);
    g_reflect_about(v, h, out);
    return 1;
  }
  return g_ds_glass(s, &d, v, rw, ru, rv, 1, out);
}

static void g_bsdf_eval(const g_surface *s, const double *l, const double *v, double *out) {
  if (s->type == 0)
    g_rp_eval(s, l, v, out);
  else if (s->type == 1)
    g_ds_eval(s, l, v, out);
  else if (s->type == 4)
    g_sg_eval(s, l, v, out);
  else
    g_ct_eval(s, l, v, out);
}
static double g_bsdf_pdf(const g_surface *s, const double *l, const double *v) {
  return s->type == 0 ? g_rp_pdf(s, l, v) : s->type == 1 ? g_ds_pdf(s, l, v) : g_ct_pdf(s, l, v);
}
/* The material's sampler on uniforms (r0, r1, r2): roughPlastic uses two.
 * Returns 0 for the reference's (-2, -2, -2). */
static int g_bsdf_sample_r(const g_surface *s, const double *v, float r0, float r1, float r2,
                           double *out) {
  if (s->type == 0) return g_rp_sample(s, v, r0, r1, out);
  if (s->type == 1) return g_ds_sample(s, v, r0, r1, r2, out);
  return g_ct_sample(s, v, r0, r1, r2, out);
}

static int g_occluded(const orc_mesh *m, const double *o, const double *d) {
  for (int64_t t = 0; t < m->n_tris; ++t) {
    const float *A = m->vertices + 3 * (size_t)m->indices[3 * t];
    const float *B = m->vertices + 3 * (size_t)m->indices[3 * t + 1];
    const float *C = m->vertices + 3 * (size_t)m->indices[3 * t + 2];
    double e1[3], e2[3], s[3], p[3], q[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = (double)B[k] - A[k];
      e2[k] = (double)C[k] - A[k];
      s[k] = o[k] - A[k];
    }
    g_cross(d, e2, p);
    const double det = g_dot(e1, p);
    if (det == 0.0) continue;
    const double u = g_dot(s, p) / det;
    if (!(u >= 0 && u <= 1)) continue;
    g_cross(s, e1, q);
    const double v = g_dot(d, q) / det;
    if (!(v >= 0 && u + v <= 1)) continue;
    if (g_dot(e2, q) / det > 1e-2) return 1;
  }
  return 0;
}

static void g_direct(const orc_mesh *m, const double *tab, int n, const float *light,
                     const double *p, const double *d, const double *ns, const double *ng,
                     const float *mat, int samples, uint32_t seed, uint32_t pixel, double *out) {
  g_surface s;
  g_surface_of(mat, ns, ng, &s);
  const double v[3] = {-d[0], -d[1], -d[2]};
  double acc[3] = {0, 0, 0};
  for (int k = 0; k < samples; ++k) {
    const uint32_t c = (uint32_t)k * 8u;
    double l[3], f[3], le[3];
    g_env_sample(tab, n, orc_rng_uniform(seed, pixel, c), orc_rng_uniform(seed, pixel, c + 1),
                 orc_rng_uniform(seed, pixel, c + 2), l);
    double pe = g_env_pdf(tab, n, l);
    if (pe > 1e-4 && !g_occluded(m, p, l)) {
      const double pb = g_bsdf_pdf(&s, l, v);
      g_bsdf_eval(&s, l, v, f);
      g_env(light, n, l, le);
      const double w = pe / fmax(pe * pe + pb * pb, 1e-8);
      for (int a = 0; a < 3; ++a) {
        const double r = f[a] * le[a] * w;
        if (!isnan(r)) acc[a] += r;
      }
    }
    const int ok = g_bsdf_sample_r(&s, v, orc_rng_uniform(seed, pixel, c + 3),
                                   orc_rng_uniform(seed, pixel, c + 4),
                                   orc_rng_uniform(seed, pixel, c + 5), l);
    if (ok) {
      const double pb = g_bsdf_pdf(&s, l, v);
      pe = g_env_pdf(tab, n, l);
      if (pe > 1e-4 && !g_occluded(m, p, l)) {
        g_bsdf_eval(&s, l, v, f);
        g_env(light, n, l, le);
        const double w = pb / fmax(pe * pe + pb * pb, 1e-8);
        for (int a = 0; a < 3; ++a) {
          const double r = f[a] * le[a] * w;
          if (!isnan(r)) acc[a] += r;
        }
      }
    }
  }
  for (int a = 0; a < 3; ++a) out[a] = acc[a] / samples;
}

void orc_gbuffer_cast_direct(const orc_mesh *m, const float *camera, int w, int h,
                             int max_bounce, const float *bounds, const float *env_faces,
                             int env_n, const float *light, int light_n, int samples,
                             uint32_t seed, double *position, double *normal, double *reflect,
                             double *kd, double *ks, double *roughness, uint8_t *env_flag,
                             int32_t *prim, double *margin, double *direct) {
  double *tab = (double *)malloc(sizeof(double) * orc_env_table_doubles(light_n));
  orc_env_tables(light, light_n, tab);
  g_cast(m, camera, w, h, max_bounce, bounds, env_faces, env_n, position, normal, reflect, kd, ks,
         roughness, env_flag, prim, margin, light, tab, light_n, samples, seed, direct);
  free(tab);
}

/* Test taps of the direct-light pieces (mat: a material row). */
void orc_ct_eval(const float *mat, const double *ns, const double *ng, const double *l,
                 const double *v, double *out) {
  g_surface s;
  g_surface_of(mat, ns, ng, &s);
  g_bsdf_eval(&s, l, v, out);
}
double orc_ct_pdf(const float *mat, const double *ns, const double *ng, const double *l,
                  const double *v) {
  g_surface s;
  g_surface_of(mat, ns, ng, &s);
  return g_bsdf_pdf(&s, l, v);
}
double orc_env_pdf(const double *tab, int n, const double *d) { return g_env_pdf(tab, n, d); }

/* ---- SH-PRT render -------------------------------------------------------
 * RenderEngine::shRender (Renderer/RenderEngine.cpp:459-486) and
 * __closesthit__SHradiance (Renderer/device_functions/devicePrograms.cu:
 * 312-501) in FP64, replaying the product's random stream: vertex colours
 * per triangle corner (stream keyed by seed, vertex and material), pixels
 * interpolate their first hit's corner colours. */

/* pbrt's real SH basis of degree 8 (Renderer/device_functions/
 * shaderUtils.h:230-337) at unit w: 81 values, index l (l + 1) + m. */
void orc_sh_eval(const double *w, double *y) {
  const int L = 8;
#define IDX(l, m) ((l) * (l) + (l) + (m))
  const double x = w[2];
  for (int k = 0; k < 81; ++k) y[k] = 0.0;
  y[IDX(0, 0)] = 1.0;
  y[IDX(1, 0)] = x;
  for (int l = 2; l <= L; ++l)
    y[IDX(l, 0)] = ((2 * l - 1) * x * y[IDX(l - 1, 0)] - (l - 1) * y[IDX(l - 2, 0)]) / l;
  double neg = -1.0, dfact = 1.0;
  const double xroot = sqrt(fmax(0.0, 1.0 - x * x));
  double xpow = xroot;
  for (int l = 1; l <= L; ++l) {
    y[IDX(l, l)] = neg * dfact * xpow;
    neg = -neg;
    dfact *= 2 * l + 1;
    xpow *= xroot;
  }
  for (int l = 2; l <= L; ++l) y[IDX(l, l - 1)] = x * (2 * l - 1) * y[IDX(l - 1, l - 1)];
  for (int l = 3; l <= L; ++l)
    for (int m = 1; m <= l - 2; ++m)
      y[IDX(l, m)] = ((2 * (l - 1) + 1) * x * y[IDX(l - 1, m)] - (l - 1 + m) * y[IDX(l - 2, m)]) /
                     (l - m);
  double sins[9], coss[9];
  const double xy = sqrt(fmax(0.0, 1.0 - x * x));
  if (xy == 0.0) {
    for (int i = 0; i <= L; ++i) {
      sins[i] = 0.0;
      coss[i] = 1.0;
    }
  } else {
    const double sp = w[1] / xy, cp = w[0] / xy;
    double si = 0.0, ci = 1.0;
    for (int i = 0; i <= L; ++i) {
      sins[i] = si;
      coss[i] = ci;
      const double os = si;
      si = si * cp + ci * sp;
      ci = ci * cp - os * sp;
    }
  }
  for (int l = 0; l <= L; ++l) {
    for (int m = -l; m < 0; ++m) {
      double inv = 1.0;
      for (int t = l + m + 1; t <= l - m; ++t) inv /= t;
      y[IDX(l, m)] = sqrt(2.0) * sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * inv) * y[IDX(l, -m)] *
                     sins[-m];
    }
    y[IDX(l, 0)] *= sqrt((2.0 * l + 1.0) / (4.0 * M_PI));
    for (int m = 1; m <= l; ++m) {
      double inv = 1.0;
      for (int t = l - m + 1; t <= l + m; ++t) inv /= t;
      y[IDX(l, m)] *= sqrt(2.0) * sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * inv) * coss[m];
    }
  }
#undef IDX
}

void orc_sh_render(const orc_mesh *m, const float *camera, int w, int h, const float *transfer,
                   const float *light_sh, int degree, int samples, uint32_t seed,
                   const float *env_faces, int env_n, double *color, int32_t *prim,
                   double *margin) {
  const int nc = (degree + 1) * (degree + 1);
  double *corner = (double *)calloc((size_t)m->n_tris * 9, sizeof(double));
  const double cam[3] = {camera[0], camera[1], camera[2]};
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t c = 0; c < m->n_tris * 3; ++c) {
    const int64_t tri = c / 3;
    const int vi = m->indices[c];
    const int mid = m->material_ids ? m->material_ids[tri] : 0;
    const float *mat = m->materials + ORC_MAT_FLOATS * (size_t)mid;
    g_surface s;
    double nrm[3], view[3];
    for (int a = 0; a < 3; ++a) {
      nrm[a] = m->normals[3 * (size_t)vi + a];
      view[a] = cam[a] - (double)m->vertices[3 * (size_t)vi + a];
    }
    g_norm(nrm);
    g_norm(view);
    g_surface_of(mat, nrm, nrm, &s);
    /* light response R_k = sum_j T[v][k][j] L[j] */
    double lr[81][3];
    for (int k = 0; k < nc; ++k)
      for (int ch = 0; ch < 3; ++ch) {
        double acc = 0.0;
        for (int j = 0; j < nc; ++j)
          acc += (double)transfer[(((size_t)vi * nc + k) * nc + j) * 3 + ch] *
                 (double)light_sh[3 * j + ch];
        lr[k][ch] = acc;
      }
    const uint32_t sd = seed ^ (0x9E3779B9u * (uint32_t)(mid + 1));
    double acc[3] = {0, 0, 0}, y[81], l[3], f[3];
    int valid = 0;
    for (int k = 0; k < samples; ++k) {
      const uint32_t q = (uint32_t)k * 4u;
      const int ok = g_bsdf_sample_r(&s, view, orc_rng_uniform(sd, (uint32_t)vi, q),
                                     orc_rng_uniform(sd, (uint32_t)vi, q + 1),
                                     orc_rng_uniform(sd, (uint32_t)vi, q + 2), l);
      if (!ok) continue;
      const double pdf = g_bsdf_pdf(&s, l, view);
      if (pdf < 1e-6) continue;
      g_bsdf_eval(&s, l, view, f);
      if (isnan(f[0]) || isnan(f[1]) || isnan(f[2])) continue;
      orc_sh_eval(l, y);
      for (int ch = 0; ch < 3; ++ch) {
        double d = 0.0;
        for (int j = 0; j < nc; ++j) d += y[j] * lr[j][ch];
        acc[ch] += f[ch] / pdf * d;
      }
      ++valid;
    }
    for (int ch = 0; ch < 3; ++ch) corner[9 * tri + 3 * (c % 3) + ch] = valid ? fabs(acc[ch] / valid) : 0.0;
  }
  /* primary rays, as orc_gbuffer_cast casts them */
  double dir[3], up[3], hz[3], vt[3];
  for (int a = 0; a < 3; ++a) {
    dir[a] = (double)camera[3 + a] - camera[a];
    up[a] = camera[6 + a];
  }
  g_norm(dir);
  g_cross(dir, up, hz);
  g_norm(hz);
  for (int a = 0; a < 3; ++a) hz[a] *= (double)w / (double)h;
  g_cross(hz, dir, vt);
  g_norm(vt);
  const double th = tan(0.5 * 0.8575560548920328);
  const double cx = w / 2.0, cy = h / 2.0, fx = 0.5 * w / th, fy = 0.5 * h / th;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t px = 0; px < (int64_t)w * h; ++px) {
    const int ix = (int)(px % w), iy = (int)(px / w);
    const double sx = (ix - cx) / fx, sy = (iy - cy) / fy;
    double d[3];
    for (int a = 0; a < 3; ++a) d[a] = dir[a] + sx * hz[a] + sy * vt[a];
    g_norm(d);
    double t, u, v, mg = 1.0, c[3] = {0, 0, 0};
    const int hit = g_closest(m, cam, d, &t, &u, &v, &mg);
    if (hit < 0) {
      if (env_faces) g_env(env_faces, env_n, d, c);
    } else {
      const double *cc = corner + 9 * (size_t)hit;
      for (int ch = 0; ch < 3; ++ch) c[ch] = (1.0 - u - v) * cc[ch] + u * cc[3 + ch] + v * cc[6 + ch];
    }
    for (int ch = 0; ch < 3; ++ch) color[4 * px + ch] = c[ch];
    color[4 * px + 3] = 1.0;
    prim[px] = hit;
    margin[px] = mg;
  }
  free(corner);
}

/* ---- SH transfer matrices ------------------------------------------------
 * RenderEngine::shComputeTMatrix (Renderer/RenderEngine.cpp:410-446) with its
 * ray generation program __raygen__ComputeTMatrix (Renderer/device_functions/
 * cameraShTmatrix.cu:87-178) and the radiance closest-hit and miss programs
 * it traces (devicePrograms.cu:47-311, 517-564), in FP64, literally: for
 * each basis map i (which is at once the view, lighting and sampling
 * environment, RenderEngine.cpp:427-428) and each vertex, `samples` paths
 * start at the vertex in uniform sphere directions (sample_sphere,
 * shaderUtils.h:151-158).  A path adds, at every non-mirror hit, one
 * environment-importance sample (shadow-tested, power-heuristic weight)
 * times the path throughput, and on leaving the scene after the first ray
 * the environment radiance with the BSDF-sample weight; it is cut by
 * Russian roulette after depth 3 and at max_bounce.  Radiance that reached
 * the vertex straight from the environment (first ray) or through one
 * mirror (first two rays) is removed; the rest, projected onto the
 * filtered SH basis 1 / (1 + 0.005 l^2 (l+1)^2) Y_k of the start direction
 * times 4 pi / samples, accumulates into T[v][k][i]; a path whose result is
 * NaN scales the vertex's row by 1 + 1 / samples instead (the reference's
 * T += T / N).  Output layout [vertex][k * nc + i][3] as RenderEngine.cpp:
 * 432-438 writes it.  Random numbers are the product's counter-based stream
 * keyed by (seed, vertex) with counters
 *   sample * 256 + {0, 1}                     start direction,
 *   sample * 256 + 16 (b + 1) + {0, 1, 2}      environment sample at trace b,
 *                               + {3, 4, 5}      BSDF light sample (max_bounce 1),
 *                               + {6, 7, 8}      continuation sample,
 *                               + 9              Russian roulette,
 * identical for every basis map, as the reference's per-vertex stream is
 * (one seed per launch, RenderEngine.cpp:427-429). */

/* Occlusion with a robustness figure: lowers *margin to the barycentric
 * distance to an edge of any triangle the ray passes close to within
 * (tmin, inf) and to the relative distance of a hit to tmin. */
static int g_occluded_m(const orc_mesh *m, const double *o, const double *d, double *margin) {
  const double tmin = 1e-2;
  int hit = 0;
  for (int64_t t = 0; t < m->n_tris; ++t) {
    const float *A = m->vertices + 3 * (size_t)m->indices[3 * t];
    const float *B = m->vertices + 3 * (size_t)m->indices[3 * t + 1];
    const float *C = m->vertices + 3 * (size_t)m->indices[3 * t + 2];
    double e1[3], e2[3], s[3], p[3], q[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = (double)B[k] - A[k];
      e2[k] = (double)C[k] - A[k];
      s[k] = o[k] - A[k];
    }
    g_cross(d, e2, p);
    const double det = g_dot(e1, p);
    if (det == 0.0) continue;
    const double u = g_dot(s, p) / det;
    g_cross(s, e1, q);
    const double v = g_dot(d, q) / det;
    const double tt = g_dot(e2, q) / det;
    if (!(tt > tmin * (1.0 - 1e-6))) continue;
    const double e = fmin(u, fmin(v, 1.0 - u - v));
    if (fabs(e) < *margin) *margin = fabs(e);
    if (fabs(tt - tmin) / tmin < *margin) *margin = fabs(tt - tmin) / tmin;
    if (e >= 0.0 && tt > tmin) hit = 1;
  }
  return hit;
}

static double g_rel_gap(double a, double b) {
  const double s = fmax(fabs(a), fabs(b));
  return s > 0 ? fabs(a - b) / s : 1.0;
}

/* Lobe-choice robustness of a BSDF sample drawn with uniform r (the first
 * uniform picks the lobe against the specular probability). */
static double g_lobe_gap(const g_surface *s, float r) {
  double spec;
  if (s->type == 1) { /* DisneyBSDF: the cumulative lobe probabilities */
    g_dlobes d;
    const double vv[3] = {s->ns[0], s->ns[1], s->ns[2]};
    g_disney_lobes(s, vv, &d);
    const double tot = d.dw + d.mw + d.gw + d.cw;
    const double c1 = d.dw / tot, c2 = c1 + d.mw / tot, c3 = c2 + d.gw / tot;
    return fmin(fabs((double)r - c1), fmin(fabs((double)r - c2), fabs((double)r - c3)));
  }
  if (s->type == 0) {
    const double ls = g_lum(s->ks), lr = g_lum(s->kd);
    if (ls + lr <= 0) return 1.0;
    spec = ls / (ls + lr);
  } else if (!g_spec_probability(s, &spec)) {
    return 1.0;
  }
  return fabs((double)r - spec);
}

static int g_bsdf_sample(const g_surface *s, const double *v, uint32_t seed, uint32_t vtx,
                         uint32_t c, double *out, double *margin) {
  const float r0 = orc_rng_uniform(seed, vtx, c), r1 = orc_rng_uniform(seed, vtx, c + 1);
  const double g = g_lobe_gap(s, r0);
  if (g < *margin) *margin = g;
  return g_bsdf_sample_r(s, v, r0, r1, orc_rng_uniform(seed, vtx, c + 2), out);
}

/* One path of basis map i from vertex position o0 (sample s); returns the
 * indirect radiance in ind[3] and its start direction in d0[3]. */
static void g_transfer_path(const orc_mesh *m, const float *faces, const double *tab, int n,
                            int max_bounce, int rr_depth, uint32_t seed, uint32_t vtx, int s,
                            const double *o0, double *d0, double *ind, double *margin) {
  const uint32_t base = (uint32_t)s * 256u;
  {
    const float z1 = orc_rng_uniform(seed, vtx, base), z2 = orc_rng_uniform(seed, vtx, base + 1);
    const double theta = 2.0 * M_PI * z1, phi = acos(1.0 - 2.0 * z2);
    d0[0] = sin(phi) * cos(theta);
    d0[1] = sin(phi) * sin(theta);
    d0[2] = cos(phi);
  }
  double o[3] = {o0[0], o0[1], o0[2]}, d[3] = {d0[0], d0[1], d0[2]};
  double color[3] = {0, 0, 0}, lc[3] = {1, 1, 1}, tent[3] = {0, 0, 0}, pdf = 0.0;
  int first_hit = 0, mirror_first = 0, env_hit = 0, done = 0, depth = 0;
  for (;;) {
    const uint32_t c = base + 16u * (uint32_t)(depth + 1);
    double t, u, v, rd[3] = {d[0], d[1], d[2]};
    g_norm(rd);
    const int hit = g_closest(m, o, d, &t, &u, &v, margin);
    if (hit < 0) {
      double rad[3];
      g_env(faces, n, d, rad);
      if (depth == 0) {
        for (int a = 0; a < 3; ++a) color[a] = rad[a];
      } else if (pdf < 0) {
        for (int a = 0; a < 3; ++a) color[a] += rad[a] * lc[a];
      } else {
        const double pe = g_env_pdf(tab, n, d);
        const double w = pdf * pdf / fmax(pdf * pdf + pe * pe, 1e-14);
        for (int a = 0; a < 3; ++a) color[a] += rad[a] * lc[a] * w;
      }
      done = 1;
      env_hit = 1;
    } else {
      const int *tri = m->indices + 3 * (size_t)hit;
      const float *A = m->vertices + 3 * (size_t)tri[0];
      const float *B = m->vertices + 3 * (size_t)tri[1];
      const float *C = m->vertices + 3 * (size_t)tri[2];
      double e1[3], e2[3], ng[3], ns[3], p[3];
      for (int a = 0; a < 3; ++a) {
        e1[a] = (double)B[a] - A[a];
        e2[a] = (double)C[a] - A[a];
      }
      g_cross(e1, e2, ng);
      g_norm(ng);
      if (m->normals) {
        const double ww = 1.0 - u - v;
        for (int a = 0; a < 3; ++a)
          ns[a] = ww * m->normals[3 * (size_t)tri[0] + a] + u * m->normals[3 * (size_t)tri[1] + a] +
                  v * m->normals[3 * (size_t)tri[2] + a];
        g_norm(ns);
      } else {
        for (int a = 0; a < 3; ++a) ns[a] = ng[a];
      }
      for (int a = 0; a < 3; ++a) p[a] = o[a] + rd[a] * t + 0.001;
      for (int a = 0; a < 3; ++a) o[a] = p[a];
      const int mid = m->material_ids ? m->material_ids[hit] : 0;
      float mbuf[ORC_MAT_FLOATS];
      const float *mr = g_material_row(m, mid, tri, u, v, mbuf);
      if ((int)mr[7] == 3) {
        const double nd = g_dot(ns, rd);
        for (int a = 0; a < 3; ++a) d[a] = rd[a] - 2.0 * nd * ns[a];
        if (!first_hit) mirror_first = 1;
      } else {
        g_surface sf;
        g_surface_of(mr, ns, ng, &sf);
        const double vv[3] = {-rd[0], -rd[1], -rd[2]};
        double l[3], f[3], rad[3];
        /* environment-importance light sample */
        g_env_sample(tab, n, orc_rng_uniform(seed, vtx, c), orc_rng_uniform(seed, vtx, c + 1),
                     orc_rng_uniform(seed, vtx, c + 2), l);
        double pe = g_env_pdf(tab, n, l);
        {
          const double g = g_rel_gap(pe, 1e-4);
          if (g < *margin) *margin = g;
        }
        if (!g_occluded_m(m, p, l, margin) && pe > 1e-4) {
          const double pb = g_bsdf_pdf(&sf, l, vv);
          g_bsdf_eval(&sf, l, vv, f);
          g_env(faces, n, l, rad);
          const double w = pe / fmax(pe * pe + pb * pb, 1e-8);
          double inc[3];
          for (int a = 0; a < 3; ++a) inc[a] = f[a] * rad[a] * lc[a] * w;
          if (!(isnan(inc[0]) || isnan(inc[1]) || isnan(inc[2])))
            for (int a = 0; a < 3; ++a) color[a] += inc[a];
        }
        /* BSDF light sample: direct mode only (devicePrograms.cu:202-247) */
        if (max_bounce == 1 && g_bsdf_sample(&sf, vv, seed, vtx, c + 3, l, margin)) {
          const double pb = g_bsdf_pdf(&sf, l, vv);
          pe = g_env_pdf(tab, n, l);
          const double g = g_rel_gap(pe, 1e-4);
          if (g < *margin) *margin = g;
          if (!g_occluded_m(m, p, l, margin) && pe > 1e-4) {
            g_bsdf_eval(&sf, l, vv, f);
            g_env(faces, n, l, rad);
            const double w = pb / fmax(pe * pe + pb * pb, 1e-8);
            double inc[3];
            for (int a = 0; a < 3; ++a) inc[a] = f[a] * rad[a] * lc[a] * w;
            if (!(isnan(inc[0]) || isnan(inc[1]) || isnan(inc[2])))
              for (int a = 0; a < 3; ++a) color[a] += inc[a];
          }
        }
        /* continuation */
        if (!g_bsdf_sample(&sf, vv, seed, vtx, c + 6, l, margin)) {
          done = 1;
        } else {
          pdf = g_bsdf_pdf(&sf, l, vv);
          const double g = g_rel_gap(pdf, 1e-4);
          if (g < *margin) *margin = g;
          if (pdf <= 1e-4) {
            done = 1;
          } else {
            g_bsdf_eval(&sf, l, vv, f);
            for (int a = 0; a < 3; ++a) {
              d[a] = l[a];
              lc[a] = lc[a] * f[a] / pdf;
            }
          }
        }
        first_hit = 1;
      }
    }
    ++depth;
    const int ok = !(isnan(color[0]) || isnan(color[1]) || isnan(color[2]));
    if (ok && env_hit && (depth == 1 || (depth == 2 && mirror_first)))
      for (int a = 0; a < 3; ++a) tent[a] = color[a];
    if (!done && depth > rr_depth) {
      const float z = orc_rng_uniform(seed, vtx, c + 9);
      const double rr = fmin(fmax(lc[0], fmax(lc[1], lc[2])), 0.95);
      const double g = g_rel_gap((double)z, rr);
      if (g < *margin) *margin = g;
      if ((double)z < rr) {
        for (int a = 0; a < 3; ++a) lc[a] /= fmax(rr, 1e-10);
      } else {
        done = 1;
      }
    }
    if (done || depth >= max_bounce) break;
  }
  for (int a = 0; a < 3; ++a) ind[a] = color[a] - tent[a];
}

void orc_sh_transfer(const orc_mesh *m, const float *basis, int n, int degree, int samples,
                     int max_bounce, int rr_depth, uint32_t seed, const int64_t *verts,
                     int64_t n_verts, double *T, double *abs_sum, double *margin) {
  const int nc = (degree + 1) * (degree + 1);
  const size_t td = orc_env_table_doubles(n), face_floats = (size_t)6 * n * n * 3;
  double *tabs = (double *)malloc(sizeof(double) * td * nc);
  for (int i = 0; i < nc; ++i) orc_env_tables(basis + face_floats * i, n, tabs + td * i);
  double filt[81];
  for (int l = 0; l <= 8; ++l)
    for (int mm = -l; mm <= l; ++mm)
      filt[l * l + l + mm] = 1.0 / (1.0 + 0.005 * l * l * (l + 1.0) * (l + 1.0));
  const int64_t jobs = n_verts * nc;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t job = 0; job < jobs; ++job) {
    const int64_t vi = job / nc;
    const int i = (int)(job % nc);
    const uint32_t vtx = (uint32_t)verts[vi];
    const double o0[3] = {m->vertices[3 * (size_t)verts[vi]], m->vertices[3 * (size_t)verts[vi] + 1],
                          m->vertices[3 * (size_t)verts[vi] + 2]};
    double *Tv = T + (size_t)vi * nc * nc * 3, *Av = abs_sum ? abs_sum + (size_t)vi * nc * nc * 3 : NULL;
    double mg = 1.0;
    for (int k = 0; k < nc; ++k)
      for (int a = 0; a < 3; ++a) {
        Tv[((size_t)k * nc + i) * 3 + a] = 0.0;
        if (Av) Av[((size_t)k * nc + i) * 3 + a] = 0.0;
      }
    for (int s = 0; s < samples; ++s) {
      double d0[3], ind[3], y[81];
      g_transfer_path(m, basis + face_floats * i, tabs + td * i, n, max_bounce, rr_depth, seed,
                      vtx, s, o0, d0, ind, &mg);
      orc_sh_eval(d0, y);
      const int bad = isnan(ind[0]) || isnan(ind[1]) || isnan(ind[2]);
      for (int k = 0; k < nc; ++k)
        for (int a = 0; a < 3; ++a) {
          const size_t e = ((size_t)k * nc + i) * 3 + a;
          if (bad) {
            Tv[e] += Tv[e] / samples;
            if (Av) Av[e] += Av[e] / samples;
          } else {
            const double x = filt[k] * y[k] * ind[a] * 4.0 * M_PI / samples;
            Tv[e] += x;
            if (Av) Av[e] += fabs(x);
          }
        }
    }
    if (margin) margin[vi * nc + i] = mg;
  }
  free(tabs);
}

/* ---- BSDF test taps over many directions ------------------------------
 * The material's sampler on n uniform triples u[n][3] (ok[i] = 0 where it
 * returns the reference's (-2, -2, -2)), and its value (with the cosine)
 * and pdf at n directions l[n][3], for view direction v; mat is a material
 * row. */
void orc_bsdf_sample_many(const float *mat, const double *ns, const double *ng, const double *v,
                          int64_t n, const float *u, double *out, int32_t *ok) {
  g_surface s;
  g_surface_of(mat, ns, ng, &s);
  for (int64_t i = 0; i < n; ++i)
    ok[i] = g_bsdf_sample_r(&s, v, u[3 * i], u[3 * i + 1], u[3 * i + 2], out + 3 * i);
}
void orc_bsdf_eval_many(const float *mat, const double *ns, const double *ng, const double *v,
                        int64_t n, const double *l, double *out) {
  g_surface s;
  g_surface_of(mat, ns, ng, &s);
  for (int64_t i = 0; i < n; ++i) g_bsdf_eval(&s, l + 3 * i, v, out + 3 * i);
}
void orc_bsdf_pdf_many(const float *mat, const double *ns, const double *ng, const double *v,
                       int64_t n, const double *l, double *out) {
  g_surface s;
  g_surface_of(mat, ns, ng, &s);
  for (int64_t i = 0; i < n; ++i) out[i] = g_bsdf_pdf(&s, l + 3 * i, v);
}
