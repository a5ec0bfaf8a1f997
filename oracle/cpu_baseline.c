/*
 * cpu_baseline.c -- FP32 restatement of the relighting math, timed on the
 * host cores as the CPU baseline (checker code; see oracle.h).
 *
 * Same math as orc_relight_pixels but in FP32 and fused per pixel: hash
 * encode and projection once, then the K wavelets through the 4-layer MLP,
 * softplus and the clamped sum.  OpenMP over pixels; compiled -O3 for
 * x86-64-v3 (AVX2 + FMA) so the inner dot products vectorise.
 */
#include <math.h>
#include <omp.h>
#include <string.h>

#include "oracle.h"

static void encode_fp32(const orc_scene *sc, const float pos[3], float *out) {
  const int L = sc->n_levels, F = sc->n_features;
  float feat[64];
  for (int l = 0; l < L; ++l) {
    uint32_t g[3];
    float fr[3];
    for (int d = 0; d < 3; ++d) {
      const float p = fmaf(pos[d], sc->scale[l], 0.5f);
      const float fl = floorf(p);
      g[d] = (uint32_t)(int)fl;
      fr[d] = p - fl;
    }
    for (int j = 0; j < F; ++j) feat[l * F + j] = 0.f;
    for (int c = 0; c < 8; ++c) {
      const uint32_t cx = g[0] + (c & 1), cy = g[1] + ((c >> 1) & 1), cz = g[2] + ((c >> 2) & 1);
      const uint32_t r = sc->resolution[l];
      const uint32_t idx = sc->dense[l] ? (cx + cy * r + cz * r * r) % sc->size[l]
                                        : (cx ^ (cy * 2654435761u) ^ (cz * 805459861u)) % sc->size[l];
      const float w = ((c & 1) ? fr[0] : 1.f - fr[0]) * ((c & 2) ? fr[1] : 1.f - fr[1]) *
                      ((c & 4) ? fr[2] : 1.f - fr[2]);
      const float *e = sc->table + ((size_t)sc->offset[l] + idx) * F;
      for (int j = 0; j < F; ++j) feat[l * F + j] += w * e[j];
    }
  }
  const int Fd = sc->feature_dim;
  for (int o = 0; o < Fd; ++o) {
    if (sc->proj) {
      float acc = 0.f;
      for (int i = 0; i < Fd; ++i) acc += sc->proj[o * Fd + i] * feat[i];
      out[o] = acc;
    } else {
      out[o] = feat[o];
    }
  }
}

/* Relights pixels [pixel_begin, pixel_begin + n_pixels) into indirect
 * (float3 per pixel) and returns the elapsed wall time in seconds. */
double orc_cpu_relight_fp32(const orc_scene *sc, const int32_t *sel, const float *light, int k,
                            const orc_gbuffer *gb, int pixel_begin, int n_pixels,
                            float *indirect, int n_threads) {
  const int Fd = sc->feature_dim, H = sc->hidden, NH = sc->n_hidden_layers;
  const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads > 0 ? n_threads : 1)
  for (int q = 0; q < n_pixels; ++q) {
    const int p = pixel_begin + q;
    float rgb[3] = {0.f, 0.f, 0.f};
    if (gb->env_flag == NULL || gb->env_flag[p]) {
      float hp[64], z[64], a[128], b[128], b2[128];
      for (int o = 0; o < H; ++o) b2[o] = 0.f;
      if (sc->aux) {
        float ax[16];
        orc_aux_inputs(gb, p, sc->aux_pad, ax);
        for (int o = 0; o < H; ++o) {
          float acc = 0.f;
          for (int i = 0; i < 16; ++i) acc += sc->waux[o * 16 + i] * ax[i];
          b2[o] = acc;
        }
      }
      encode_fp32(sc, gb->position + 4 * (size_t)p, hp);
      for (int kk = 0; kk < k; ++kk) {
        const int c = sel[kk];
        for (int f = 0; f < Fd; ++f) z[f] = hp[f] * sc->wfeat[(size_t)f * sc->n_coef + c];
        const float *W = sc->layers;
        for (int o = 0; o < H; ++o) {
          float acc = 0.f;
          for (int i = 0; i < Fd; ++i) acc += W[o * Fd + i] * z[i];
          a[o] = acc > 0.f ? acc : 0.f;
        }
        W += (size_t)H * Fd;
        for (int l = 1; l < NH; ++l) {
          for (int o = 0; o < H; ++o) {
            float acc = l == 1 ? b2[o] : 0.f;
            for (int i = 0; i < H; ++i) acc += W[o * H + i] * a[i];
            b[o] = acc > 0.f ? acc : 0.f;
          }
          memcpy(a, b, sizeof(float) * (size_t)H);
          W += (size_t)H * H;
        }
        for (int ch = 0; ch < 3; ++ch) {
          float acc = 0.f;
          for (int i = 0; i < H; ++i) acc += W[ch * H + i] * a[i];
          const float t = sc->softplus ? (acc > 0.f ? acc + log1pf(expf(-acc)) : log1pf(expf(acc)))
                                       : acc;
          const float term = light[3 * kk + ch] * t;
          rgb[ch] += term > 0.f ? term : 0.f;
        }
      }
    }
    for (int ch = 0; ch < 3; ++ch) indirect[3 * (size_t)q + ch] = rgb[ch];
  }
  return omp_get_wtime() - t0;
}
