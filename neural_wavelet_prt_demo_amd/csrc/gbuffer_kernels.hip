// G-buffer producer: one primary ray per pixel through a two-child BVH in
// device memory, writing the first-hit encodings the relighting path reads.
// It stands in for the first-hit part of the reference's OptiX pass:
// - ray generation: __raygen__fovCamera (Renderer/device_functions/
//   camera.cu:82-131) with the pixel jitter at its mean, and the
//   encodings it stores (camera.cu:233-270);
// - closest hit: __closesthit__radiance (Renderer/device_functions/
//   devicePrograms.cu:47-311): barycentric normal, hit point
//   p = o + d t + 0.001, mirror reflection (material type 3), the
//   material's first-hit kd / ks / roughness;
// - miss: __miss__radiance (devicePrograms.cu:517-564): normal and reflect
//   = -d, position -1e5, kd = view environment radiance on the first ray;
// - the cube-face lookup EnvDirecToUV (Renderer/device_functions/
//   cubeMap.cu:21-92) with bilinear filtering and wrap addressing
//   (Renderer/textures.cpp:31-36).
// The ray-triangle test is Moller-Trumbore, double-sided (OptiX builds the
// GAS with no culling flags), t in (0.01, best]: the reference's tmin is
// 1e-2 (camera.cu:57).  Equal distances resolve to the lower triangle
// index, so the result does not depend on the traversal order.
#include <hip/hip_runtime.h>

#include <climits>

#include "kernels.hpp"
#include "srgb_table.hpp"

namespace nwprt {
namespace {

constexpr float kTmin = 1e-2f;
constexpr float kTmax = 1e20f;
// Box exits are widened by 2 gamma_3 (3 roundings of (lo - o) * (1 / d)),
// so rounding never culls a box the ray actually crosses.
constexpr float kExitScale = 1.0000004f;

__device__ inline float3 v3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ inline float3 v3(const float4 &a) { return make_float3(a.x, a.y, a.z); }
__device__ inline float3 v3(const float *a) { return make_float3(a[0], a[1], a[2]); }
__device__ inline float3 sub(float3 a, float3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline float3 add(float3 a, float3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline float3 scale(float3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ inline float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float3 cross(float3 a, float3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline float3 normalize(float3 a) { return scale(a, 1.0f / sqrtf(dot(a, a))); }
// reflect(i, n) = i - 2 (n . i) n
__device__ inline float3 reflect(float3 i, float3 n) { return sub(i, scale(n, 2.0f * dot(n, i))); }

struct Hit {
  float t, u, v;
  int prim, slot;  // original triangle index, position in the leaf-ordered array
};

__device__ inline bool box_entry(float3 lo, float3 hi, float3 o, float3 inv, float tmax,
                                 float &tn) {
  const float tx0 = (lo.x - o.x) * inv.x, tx1 = (hi.x - o.x) * inv.x;
  const float ty0 = (lo.y - o.y) * inv.y, ty1 = (hi.y - o.y) * inv.y;
  const float tz0 = (lo.z - o.z) * inv.z, tz1 = (hi.z - o.z) * inv.z;
  tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), kTmin));
  const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
  return tn <= tf * kExitScale;
}

__device__ inline void intersect_leaf(const BvhTri *__restrict__ tris, int ref, float3 o,
                                      float3 d, Hit &h) {
  const int first = ref >> 3, count = (ref & 7) + 1;
  for (int i = 0; i < count; ++i) {
    const float4 *r = reinterpret_cast<const float4 *>(tris + first + i);
    const float4 a = r[0], b = r[1], c = r[2];
    const float3 e1 = v3(b), e2 = v3(c);
    const float3 p = cross(d, e2);
    const float det = dot(e1, p);
    if (det == 0.0f) continue;
    const float inv = 1.0f / det;
    const float3 s = sub(o, v3(a));
    const float u = dot(s, p) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) continue;
    const float3 q = cross(s, e1);
    const float v = dot(d, q) * inv;
    if (!(v >= 0.0f && u + v <= 1.0f)) continue;
    const float t = dot(e2, q) * inv;
    const int prim = __float_as_int(a.w);
    if (t > kTmin && (t < h.t || (t == h.t && prim < h.prim))) h = Hit{t, u, v, prim, first + i};
  }
}

// Closest hit along o + t d, t in (kTmin, kTmax).  stack points at this
// lane's column of a [kBvhMaxDepth][64] LDS array.
__device__ inline Hit trace(const BvhNode *__restrict__ nodes, const BvhTri *__restrict__ tris,
                            float3 o, float3 d, int *stack) {
  Hit h{kTmax, 0.0f, 0.0f, INT_MAX, 0};
  auto safe_inv = [](float x) { return 1.0f / (fabsf(x) > 1e-30f ? x : copysignf(1e-30f, x)); };
  const float3 inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
  int node = 0, sp = 0;
  for (;;) {
    const float4 *nr = reinterpret_cast<const float4 *>(nodes + node);
    const float4 n0 = nr[0], n1 = nr[1], n2 = nr[2], n3 = nr[3];
    int c0 = __float_as_int(n0.w), c1 = __float_as_int(n1.w);
    float t0, t1;
    bool h0 = box_entry(v3(n0), v3(n1), o, inv, h.t, t0);
    bool h1 = box_entry(v3(n2), v3(n3), o, inv, h.t, t1);
    if (h0 && c0 < 0) {
      intersect_leaf(tris, ~c0, o, d, h);
      h0 = false;
    }
    if (h1 && c1 < 0) {
      intersect_leaf(tris, ~c1, o, d, h);
      h1 = false;
    }
    if (h0 && h1) {
      if (t1 < t0) {
        const int c = c0;
        c0 = c1;
        c1 = c;
      }
      // The host bounds the tree at kBvhMaxDepth internal levels and a path
      // pushes at most one entry per level, so sp stays below the bound.
      if (sp < kBvhMaxDepth) stack[64 * sp++] = c1;
      node = c0;
    } else if (h0) {
      node = c0;
    } else if (h1) {
      node = c1;
    } else {
      if (sp == 0) break;
      node = stack[64 * --sp];
    }
  }
  return h;
}

// Radiance of a cube map in direction d: face choice and (u, v) as
// EnvDirecToUV (later tests win on ties), then a bilinear, wrapped fetch of
// the [6][n][n][3] faces with texel centres at (i + 0.5) / n.
__device__ float3 env_radiance(const float *__restrict__ faces, int n, float3 d) {
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  float m = ax, uc = d.z, vc = d.y;
  int f = 0;
  if (d.x > 0.0f && ax >= ay && ax >= az) { m = ax; uc = d.z; vc = d.y; f = 0; }
  if (!(d.x > 0.0f) && ax >= ay && ax >= az) { m = ax; uc = -d.z; vc = d.y; f = 1; }
  if (d.y > 0.0f && ay >= ax && ay >= az) { m = ay; uc = d.x; vc = d.z; f = 2; }
  if (!(d.y > 0.0f) && ay >= ax && ay >= az) { m = ay; uc = d.x; vc = -d.z; f = 3; }
  if (!(d.z > 0.0f) && az >= ax && az >= ay) { m = az; uc = d.x; vc = d.y; f = 4; }
  if (d.z > 0.0f && az >= ax && az >= ay) { m = az; uc = -d.x; vc = d.y; f = 5; }
  if (!(m > 0.0f)) return v3(0.0f, 0.0f, 0.0f);
  const float u = 0.5f * (uc / m + 1.0f), v = 0.5f * (vc / m + 1.0f);
  const float x = u * (float)n - 0.5f, y = v * (float)n - 0.5f;
  const float fx = floorf(x), fy = floorf(y);
  const float wx = x - fx, wy = y - fy;
  int x0 = (int)fx % n, y0 = (int)fy % n;
  if (x0 < 0) x0 += n;
  if (y0 < 0) y0 += n;
  const int x1 = x0 + 1 == n ? 0 : x0 + 1, y1 = y0 + 1 == n ? 0 : y0 + 1;
  const float *face = faces + (size_t)f * n * n * 3;
  const float *p00 = face + ((size_t)y0 * n + x0) * 3, *p01 = face + ((size_t)y0 * n + x1) * 3;
  const float *p10 = face + ((size_t)y1 * n + x0) * 3, *p11 = face + ((size_t)y1 * n + x1) * 3;
  float r[3];
  for (int c = 0; c < 3; ++c) {
    const float top = p00[c] + wx * (p01[c] - p00[c]);
    const float bot = p10[c] + wx * (p11[c] - p10[c]);
    r[c] = top + wy * (bot - top);
  }
  return v3(r[0], r[1], r[2]);
}

__device__ inline float4 enc(float3 v) {
  return make_float4(v.x * 0.5f + 0.5f, v.y * 0.5f + 0.5f, v.z * 0.5f + 0.5f, 1.0f);
}

// Any hit along o + t d, t in (kTmin, kTmax): the shadow-ray query
// (OPTIX_RAY_FLAG_TERMINATE_ON_FIRST_HIT, devicePrograms.cu:176-196).
__device__ inline bool occluded(const BvhNode *__restrict__ nodes, const BvhTri *__restrict__ tris,
                                float3 o, float3 d, int *stack) {
  auto safe_inv = [](float x) { return 1.0f / (fabsf(x) > 1e-30f ? x : copysignf(1e-30f, x)); };
  const float3 inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
  int node = 0, sp = 0;
  for (;;) {
    const float4 *nr = reinterpret_cast<const float4 *>(nodes + node);
    const float4 n0 = nr[0], n1 = nr[1], n2 = nr[2], n3 = nr[3];
    const int c0 = __float_as_int(n0.w), c1 = __float_as_int(n1.w);
    float t0, t1;
    bool h0 = box_entry(v3(n0), v3(n1), o, inv, kTmax, t0);
    bool h1 = box_entry(v3(n2), v3(n3), o, inv, kTmax, t1);
    Hit h{kTmax, 0.0f, 0.0f, INT_MAX, 0};
    if (h0 && c0 < 0) {
      intersect_leaf(tris, ~c0, o, d, h);
      h0 = false;
    }
    if (h1 && c1 < 0) {
      intersect_leaf(tris, ~c1, o, d, h);
      h1 = false;
    }
    if (h.prim != INT_MAX) return true;
    if (h0 && h1) {
      if (sp < kBvhMaxDepth) stack[64 * sp++] = c1;
      node = c0;
    } else if (h0) {
      node = c0;
    } else if (h1) {
      node = c1;
    } else {
      if (sp == 0) return false;
      node = stack[64 * --sp];
    }
  }
}

// ------------------------------------------------------------ direct light
// The closest-hit program's direct term in wavelet mode (maximumBounce 1,
// devicePrograms.cu:165-280): per light sample, one environment-importance
// sample and one BSDF sample, each shadow-tested and weighted by the power
// heuristic, for the cookTorrance (cookTorrance.cu:26-195), roughPlastic
// and simple_ggx BSDFs.  Random
// numbers come from a counter-based generator (SplitMix64's finaliser over
// seed, pixel and a per-sample counter) so a host restatement can replay
// every sample.

__device__ inline float rng_uniform(uint32_t seed, uint32_t pixel, uint32_t counter) {
  uint64_t z = (((uint64_t)seed << 32) | pixel) + 0x9E3779B97F4A7C15ull * ((uint64_t)counter + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (float)(uint32_t)(z >> 40) * 5.9604644775390625e-8f;  // 2^-24
}

// Separately rounded products and sums (no contraction), as the oracle
// evaluates the reference's FP32 luminance.
__device__ inline float luminance(float3 c) {
#pragma clang fp contract(off)
  return c.x * 0.212671f + c.y * 0.715160f + c.z * 0.072169f;
}

// Orthonormal frame around a unit normal (Frisvad 2012), as the reference's
// to_local / to_world use it (shaderUtils.h:16-46).
__device__ inline void frame(float3 n, float3 &t, float3 &b) {
  if (n.z < -1.0f + 1e-6f) {
    t = v3(0.0f, -1.0f, 0.0f);
    b = v3(-1.0f, 0.0f, 0.0f);
  } else {
    const float a = 1.0f / (1.0f + n.z), c = -n.x * n.y * a;
    t = v3(1.0f - n.x * n.x * a, c, -n.x);
    b = v3(c, 1.0f - n.y * n.y * a, -n.y);
  }
}
__device__ inline float3 to_local(float3 d, float3 n) {
  n = normalize(n);
  float3 t, b;
  frame(n, t, b);
  return v3(dot(d, t), dot(d, b), dot(d, n));
}
__device__ inline float3 to_world(float3 d, float3 n) {
  n = normalize(n);
  float3 t, b;
  frame(n, t, b);
  if (!(n.z < -1.0f + 1e-6f)) {
    t = normalize(t);
    b = normalize(b);
  }
  return add(add(scale(t, d.x), scale(b, d.y)), scale(n, d.z));
}

// Trowbridge-Reitz (GTR2) distribution and its Smith masking term.
__device__ inline float gtr2(float3 h, float ax, float ay) {
  const float x = h.x / ax, y = h.y / ay, l2 = x * x + y * y + h.z * h.z;
  return 1.0f / (3.14159265358979f * ax * ay * l2 * l2);
}
__device__ inline float smith_gtr2(float3 v, float ax, float ay) {
  const float lam = (-1.0f + sqrtf(1.0f + (v.x * v.x * ax * ax + v.y * v.y * ay * ay) / (v.z * v.z))) * 0.5f;
  return 1.0f / (1.0f + lam);
}

struct Surface {
  float3 kd, ks;
  float rough;
  float3 ns, ng;  // unit shading and geometric normals
  int type;       // 0 roughPlastic, 1 DisneyBSDF, 2 cookTorrance, 4 simple_ggx
  // DisneyBSDF parameters; eta also for roughPlastic.
  float st, metallic, subsurface, specular, spec_tint, aniso, sheen, sheen_tint, clearcoat,
      cc_gloss, eta;
};

// One texel of a material texture as the reference's CUDA texture object
// reads it: b / 255, the colour channels through the sRGB decode.
__device__ inline float4 texel(const TexRef &t, int i, int j) {
  // The texel pointer comes from the mesh's texture table in memory: typed
  // as global so the fetch is a global, not a flat, load.
  typedef const __attribute__((address_space(1))) uint32_t *global_u32;
  const uint32_t q = ((global_u32)t.rgba)[(size_t)j * t.w + i];
  return make_float4(kSrgbToLinear[q & 0xffu], kSrgbToLinear[(q >> 8) & 0xffu],
                     kSrgbToLinear[(q >> 16) & 0xffu], (float)(q >> 24) * (1.0f / 255.0f));
}

// tex2D with normalized coordinates, wrap addressing and bilinear filtering
// (Renderer/textures.cpp:8-47): the coordinate's fractional part times the
// size, minus half a texel, between the centres of the four neighbours.
__device__ inline float4 tex_sample(const TexRef &t, float u, float v) {
  float fu = u - floorf(u), fv = v - floorf(v);
  if (!(fu >= 0.0f && fu < 1.0f)) fu = 0.0f;  // NaN or inf coordinates
  if (!(fv >= 0.0f && fv < 1.0f)) fv = 0.0f;
  const float x = fu * (float)t.w - 0.5f, y = fv * (float)t.h - 0.5f;
  const float x0f = floorf(x), y0f = floorf(y);
  const float ax = x - x0f, ay = y - y0f;
  int x0 = (int)x0f, y0 = (int)y0f;  // in [-1, w - 1] and [-1, h - 1]
  if (x0 < 0) x0 += t.w;
  if (y0 < 0) y0 += t.h;
  const int x1 = x0 + 1 == t.w ? 0 : x0 + 1, y1 = y0 + 1 == t.h ? 0 : y0 + 1;
  const float4 a = texel(t, x0, y0), b = texel(t, x1, y0), c = texel(t, x0, y1),
               d = texel(t, x1, y1);
  const float w00 = (1.0f - ax) * (1.0f - ay), w10 = ax * (1.0f - ay), w01 = (1.0f - ax) * ay,
              w11 = ax * ay;
  return make_float4(w00 * a.x + w10 * b.x + w01 * c.x + w11 * d.x,
                     w00 * a.y + w10 * b.y + w01 * c.y + w11 * d.y,
                     w00 * a.z + w10 * b.z + w01 * c.z + w11 * d.z,
                     w00 * a.w + w10 * b.w + w01 * c.w + w11 * d.w);
}

// The texture coordinates of a hit: (1 - u - v) t0 + u t1 + v t2
// (devicePrograms.cu:66-71).
__device__ inline float2 hit_texcoord(const float *tc, const int32_t *indices, int prim, float u,
                                      float v) {
  const int *tri = indices + 3 * (size_t)prim;
  const float w = 1.0f - u - v;
  const float2 t0 = reinterpret_cast<const float2 *>(tc)[tri[0]];
  const float2 t1 = reinterpret_cast<const float2 *>(tc)[tri[1]];
  const float2 t2 = reinterpret_cast<const float2 *>(tc)[tri[2]];
  return make_float2(w * t0.x + u * t1.x + v * t2.x, w * t0.y + u * t1.y + v * t2.y);
}

// The material with its textured fields read at tc (devicePrograms.cu:
// 121-147): kd and ks take the texel's RGB, the others its first channel.
__device__ inline MaterialRec textured(MaterialRec m, const TexRef *mt, float2 tc) {
#pragma unroll 1
  for (int f = 0; f < kTexFields; ++f) {
    const TexRef t = mt[f];
    if (!t.rgba) continue;
    const float4 x = tex_sample(t, tc.x, tc.y);
    switch (f) {
      case NWPRT_TEX_KD: m.kd[0] = x.x, m.kd[1] = x.y, m.kd[2] = x.z; break;
      case NWPRT_TEX_KS: m.ks[0] = x.x, m.ks[1] = x.y, m.ks[2] = x.z; break;
      case NWPRT_TEX_ROUGHNESS: m.roughness = x.x; break;
      case NWPRT_TEX_SPECULAR_TRANSMISSION: m.specular_transmission = x.x; break;
      case NWPRT_TEX_METALLIC: m.metallic = x.x; break;
      case NWPRT_TEX_SUBSURFACE: m.subsurface = x.x; break;
      case NWPRT_TEX_SPECULAR: m.specular = x.x; break;
      case NWPRT_TEX_SPECULAR_TINT: m.specular_tint = x.x; break;
      case NWPRT_TEX_ANISOTROPIC: m.anisotropic = x.x; break;
      case NWPRT_TEX_SHEEN: m.sheen = x.x; break;
      case NWPRT_TEX_SHEEN_TINT: m.sheen_tint = x.x; break;
      case NWPRT_TEX_CLEARCOAT: m.clearcoat = x.x; break;
      case NWPRT_TEX_CLEARCOAT_GLOSS: m.clearcoat_gloss = x.x; break;
      default: m.eta = x.x; break;
    }
  }
  return m;
}

__device__ inline Surface make_surface(const MaterialRec &m, float3 ns, float3 ng) {
  return Surface{v3(m.kd), v3(m.ks), m.roughness, ns, ng, m.type, m.specular_transmission,
                 m.metallic, m.subsurface, m.specular, m.specular_tint, m.anisotropic, m.sheen,
                 m.sheen_tint, m.clearcoat, m.clearcoat_gloss, m.eta};
}

__device__ float3 ct_eval(const Surface &s, float3 l, float3 v) {
  if (dot(s.ng, v) < 0.0f || dot(s.ng, l) < 0.0f) return v3(0.0f, 0.0f, 0.0f);
  const float3 n = dot(s.ns, v) < 0.0f ? scale(s.ns, -1.0f) : s.ns;
  if (luminance(s.ks) + luminance(s.kd) <= 0.0f) return v3(0.0f, 0.0f, 0.0f);
  float3 h = normalize(add(v, l));
  const float n_in = dot(n, v), n_out = dot(n, l);
  const float alpha = fmaxf(0.0025f, s.rough * s.rough);
  if (dot(h, n) < 0.0f) h = scale(h, -1.0f);
  const float h_out = dot(h, l);
  const float sch = powf(fmaxf(1.0f - h_out, 0.0f), 5.0f);
  const float d = gtr2(to_local(h, n), alpha, alpha);
  const float g = smith_gtr2(to_local(v, n), alpha, alpha) * smith_gtr2(to_local(l, n), alpha, alpha);
  const float k = d * g / (4.0f * n_in * n_out);
  const float so = powf(1.0f - n_out, 5.0f), si = powf(1.0f - n_in, 5.0f);
  const float base = 1.05f * (1.0f - so) * (1.0f - si) / 3.14159265358979f;
  float r[3];
  const float kd[3] = {s.kd.x, s.kd.y, s.kd.z}, ks[3] = {s.ks.x, s.ks.y, s.ks.z};
  for (int c = 0; c < 3; ++c) {
    const float f = ks[c] + (1.0f - ks[c]) * sch;
    r[c] = ((1.0f - ks[c]) * base * kd[c] + f * k) * n_out;
  }
  return v3(r[0], r[1], r[2]);
}

// Specular-lobe probability: by luminance (cookTorrance), or
// clamp(1 - roughness, 0.05, 0.75) for simple_ggx (simple_ggx.cu:54-55).
__device__ inline bool spec_probability(const Surface &s, float &spec) {
  if (s.type == 4) {
    spec = fminf(fmaxf(1.0f - fminf(fmaxf(s.rough, 0.05f), 1.0f), 0.05f), 0.75f);
    return true;
  }
  const float ls = luminance(s.ks), lr = luminance(s.kd);
  if (ls + lr <= 0.0f) return false;
  spec = ls / (ls + lr);
  return true;
}

__device__ float ct_pdf(const Surface &s, float3 l, float3 v) {
  if (dot(s.ng, v) < 0.0f || dot(s.ng, l) < 0.0f) return 0.0f;
  const float3 n = dot(s.ns, v) < 0.0f ? scale(s.ns, -1.0f) : s.ns;
  const float3 h = normalize(add(v, l));
  const float n_h = dot(n, h), n_in = dot(n, v), n_out = dot(n, l);
  if (n_out <= 0.0f || n_h <= 0.0f) return 0.0f;
  float spec;
  if (!spec_probability(s, spec)) return 0.0f;
  const float r = fminf(fmaxf(s.rough, 0.05f), 1.0f), a = r * r;
  const float diff = 1.0f - spec;
  spec *= smith_gtr2(to_local(v, n), a, a) * gtr2(to_local(h, n), a, a) / (4.0f * n_in);
  return spec + diff * n_out / 3.14159265358979f;
}

// Lobe choice (spec_probability), then a cosine-weighted direction or
// Heitz's (2018) sample of the visible GGX normals; returns x = -2 for a
// black BSDF.  Shared by cookTorrance and simple_ggx (simple_ggx.cu:129-177).
__device__ float3 ct_sample(const Surface &s, float3 v, float rw, float ru, float rv) {
  const float3 n = dot(s.ns, v) < 0.0f ? scale(s.ns, -1.0f) : s.ns;
  float spec;
  if (!spec_probability(s, spec)) return v3(-2.0f, -2.0f, -2.0f);
  if (rw > spec) {
    const float phi = 2.0f * 3.14159265358979f * ru;
    const float tmp = sqrtf(fminf(fmaxf(1.0f - rv, 0.0f), 1.0f));
    return to_world(v3(cosf(phi) * tmp, sinf(phi) * tmp, sqrtf(fminf(fmaxf(rv, 0.0f), 1.0f))), n);
  }
  float3 li = to_local(v, n);
  const bool below = li.z < 0.0f;
  if (below) li = scale(li, -1.0f);
  const float r = fminf(fmaxf(s.rough, 0.05f), 1.0f), a = r * r;
  const float3 hemi = normalize(v3(a * li.x, a * li.y, li.z));
  const float rr = sqrtf(ru), phi = 2.0f * 3.14159265358979f * rv;
  const float t1 = rr * cosf(phi);
  float t2 = rr * sinf(phi);
  const float sm = (1.0f + hemi.z) * 0.5f;
  t2 = (1.0f - sm) * sqrtf(1.0f - t1 * t1) + sm * t2;
  const float3 disk = v3(t1, t2, sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2)));
  const float3 hn = to_world(disk, hemi);
  const float3 micro = normalize(v3(a * hn.x, a * hn.y, fmaxf(0.0f, hn.z)));
  const float3 hw = normalize(to_world(micro, n));
  const float3 out = normalize(add(scale(v, -1.0f), scale(hw, 2.0f * dot(v, hw))));
  return below ? scale(out, -1.0f) : out;
}

// roughPlastic (roughPlastic.cu:25-174): a GGX coat over a diffuse base,
// both weighted by the dielectric Fresnel term at the material's index of
// refraction eta (Material.h:55, default 1.49 / 1.000277, :94-97).

__device__ inline float fresnel_dielectric(float n_dot_i, float eta) {
  const float t2 = 1.0f - (1.0f - n_dot_i * n_dot_i) / (eta * eta);
  if (t2 < 0.0f) return 1.0f;
  const float nt = sqrtf(t2), ni = fabsf(n_dot_i);
  const float rs = (ni - eta * nt) / (ni + eta * nt), rp = (eta * ni - nt) / (eta * ni + nt);
  return (rs * rs + rp * rp) * 0.5f;
}

__device__ float3 rp_eval(const Surface &s, float3 l, float3 v) {
  // (the reference computes, then ignores, the geometric-normal test here)
  const float3 n = dot(s.ns, v) < 0.0f ? scale(s.ns, -1.0f) : s.ns;
  const float3 h = normalize(add(v, l));
  const float n_h = dot(n, h), n_in = dot(n, v), n_out = dot(n, l);
  if (n_out <= 0.0f || n_h <= 0.0f) return v3(0.0f, 0.0f, 0.0f);
  const float r = fminf(fmaxf(s.rough, 0.01f), 1.0f), a = r * r;
  const float fo = fresnel_dielectric(dot(h, l), s.eta);
  const float d = gtr2(to_local(h, n), a, a);
  const float g = smith_gtr2(to_local(v, n), a, a) * smith_gtr2(to_local(l, n), a, a);
  const float k = g * fo * d / (4.0f * n_in * n_out);
  const float fi = fresnel_dielectric(dot(h, v), s.eta);
  const float base = (1.0f - fo) * (1.0f - fi) / 3.14159265358979f;
  return v3((s.kd.x * base + s.ks.x * k) * n_out, (s.kd.y * base + s.ks.y * k) * n_out,
            (s.kd.z * base + s.ks.z * k) * n_out);
}

__device__ float rp_pdf(const Surface &s, float3 l, float3 v) {
  if (dot(s.ng, v) < 0.0f || dot(s.ng, l) < 0.0f) return 0.0f;
  const float3 n = dot(s.ns, v) < 0.0f ? scale(s.ns, -1.0f) : s.ns;
  const float3 h = normalize(add(v, l));
  const float n_h = dot(n, h), n_in = dot(n, v), n_out = dot(n, l);
  if (n_out <= 0.0f || n_h <= 0.0f) return 0.0f;
  const float ls = luminance(s.ks), lr = luminance(s.kd);
  if (ls + lr <= 0.0f) return 0.0f;
  const float r = fminf(fmaxf(s.rough, 0.01f), 1.0f), a = r * r;
  float spec = ls / (ls + lr);
  const float diff = 1.0f - spec;
  spec *= smith_gtr2(to_local(v, n), a, a) * gtr2(to_local(h, n), a, a) / (4.0f * n_in);
  return spec + diff * n_out / 3.14159265358979f;
}

// Two uniforms: z1 picks the lobe and is then reused for the direction, as
// the reference does.
__device__ float3 rp_sample(const Surface &s, float3 v, float z1, float z2) {
  if (dot(s.ng, v) < 0.0f) return v3(-2.0f, -2.0f, -2.0f);
  const float3 n = dot(s.ns, v) < 0.0f ? scale(s.ns, -1.0f) : s.ns;
  const float ls = luminance(s.ks), lr = luminance(s.kd);
  if (ls + lr <= 0.0f) return v3(-2.0f, -2.0f, -2.0f);
  const float spec = ls / (ls + lr);
  if (z1 > spec) {
    const float phi = 2.0f * 3.14159265358979f * z1;
    const float tmp = sqrtf(fminf(fmaxf(1.0f - z2, 0.0f), 1.0f));
    return to_world(v3(cosf(phi) * tmp, sinf(phi) * tmp, sqrtf(fminf(fmaxf(z2, 0.0f), 1.0f))), n);
  }
  const float3 li = to_local(v, n);
  const float r = fminf(fmaxf(s.rough, 0.01f), 1.0f), a = r * r;
  const float3 hemi = normalize(v3(a * li.x, a * li.y, li.z));
  const float rr = sqrtf(z1), phi = 2.0f * 3.14159265358979f * z2;
  const float t1 = rr * cosf(phi);
  float t2 = rr * sinf(phi);
  const float sm = (1.0f + hemi.z) * 0.5f;
  t2 = (1.0f - sm) * sqrtf(1.0f - t1 * t1) + sm * t2;
  const float3 disk = v3(t1, t2, sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2)));
  const float3 hn = to_world(disk, hemi);
  const float3 micro = normalize(v3(a * hn.x, a * hn.y, fmaxf(0.0f, hn.z)));
  const float3 hw = normalize(to_world(micro, n));
  return normalize(add(scale(v, -1.0f), scale(hw, 2.0f * dot(v, hw))));
}

// simple_ggx's value (simple_ggx.cu:74-126): a Schlick-Fresnel GGX
// specular term with F0 = 0.04 and Schlick-GGX geometry k = (a + 2r + 1) / 8,
// cosines clamped to [1e-6, 1], plus a Lambertian kd / pi.
__device__ float3 sg_eval(const Surface &s, float3 l, float3 v) {
  if (dot(s.ng, v) < 0.0f || dot(s.ng, l) < 0.0f) return v3(0.0f, 0.0f, 0.0f);
  const float3 n = dot(s.ns, v) < 0.0f ? scale(s.ns, -1.0f) : s.ns;
  const float3 h = normalize(add(v, l));
  auto c01 = [](float x) { return fminf(fmaxf(x, 1e-6f), 1.0f); };
  const float n_l = c01(dot(n, l)), n_v = c01(dot(n, v)), n_h = c01(dot(n, h)), v_h = c01(dot(h, v));
  const float a = s.rough * s.rough, a2 = a * a;
  const float k = (float)(((double)a + 2.0 * s.rough + 1.0) / 8.0);
  const float fmi = (float)((-5.55473 * (double)v_h - 6.98316) * (double)v_h);
  const float frac = (0.04f + 0.96f * powf(2.0f, fmi)) * a2;
  const float nom0 = n_h * n_h * (a2 - 1.0f) + 1.0f;
  const float nom1 = n_v * (1.0f - k) + k, nom2 = n_l * (1.0f - k) + k;
  const float nom = fminf(fmaxf(4.0f * 3.14159265358979f * nom0 * nom0 * nom1 * nom2, 1e-6f),
                          4.0f * 3.14159265358979f);
  const float sp = frac / nom * n_l, dif = n_l / 3.14159265358979f;
  return v3(sp + s.kd.x * dif, sp + s.kd.y * dif, sp + s.kd.z * dif);
}

// DisneyBSDF (disneyBSDF.cu:44-492): diffuse with the subsurface blend,
// sheen, anisotropic GTR2 metal, GTR1 clearcoat and a rough dielectric
// glass lobe (reflection and refraction), mixed by the weights of the
// reference, with its pdf and sampler as they are written -- including the
// diffuse pdf's cosine of the view direction and the lobe probabilities
// left unscaled in the refraction case, which the estimators' expectations
// depend on.
__device__ inline float gtr1(float n_dot_h, float a) {
  const float a2 = a * a, t = 1.0f + (a2 - 1.0f) * n_dot_h * n_dot_h;
  return (a2 - 1.0f) / (3.14159265358979f * logf(a2) * t);
}
__device__ inline float schlick5(float c) { return powf(fmaxf(1.0f - c, 0.0f), 5.0f); }

struct DisneyLobes {
  float3 n;  // shading normal flipped to the view's geometric side
  float ax, ay, ac, cc;
  float dw, mw, gw, cw;  // lobe weights
};
__device__ inline DisneyLobes disney_lobes(const Surface &s, float3 v) {
  DisneyLobes d;
  d.n = dot(s.ns, v) * dot(s.ng, v) < 0.0f ? scale(s.ns, -1.0f) : s.ns;
  d.cc = 0.25f * s.clearcoat;
  const float aspect = sqrtf(1.0f - s.aniso * 0.9f);
  d.ax = fmaxf(0.0001f, s.rough * s.rough / aspect);
  d.ay = fmaxf(0.0001f, s.rough * s.rough * aspect);
  d.ac = (1.0f - s.cc_gloss) * 0.1f + s.cc_gloss * 0.001f;
  d.dw = (1.0f - s.metallic) * (1.0f - s.st);
  d.mw = 1.0f - s.st * (1.0f - s.metallic);
  d.gw = (1.0f - s.metallic) * s.st;
  d.cw = d.cc;
  return d;
}
__device__ inline float3 tint_of(float3 c) {
  const float l = luminance(c);
  return l > 0.0f ? scale(c, 1.0f / l) : v3(1.0f, 1.0f, 1.0f);
}

__device__ float3 ds_eval(const Surface &s, float3 l, float3 v) {
  const DisneyLobes d = disney_lobes(s, v);
  const float3 n = d.n;
  const bool reflect = dot(s.ng, l) * dot(s.ng, v) > 0.0f;
  if (reflect) {
    const float n_in = dot(n, v), n_out = dot(n, l);
    float3 h = normalize(add(v, l));
    if (dot(h, n) < 0.0f) h = scale(h, -1.0f);
    const float n_h = dot(s.ns, h), h_in = dot(h, v), h_out = dot(h, l);
    float3 r = v3(0.0f, 0.0f, 0.0f);
    if (dot(s.ng, v) >= 0.0f && dot(s.ng, l) >= 0.0f && n_out > 0.0f) {
      const float fd90 = 0.5f + 2.0f * s.rough * h_out * h_out;
      const float so = powf(1.0f - n_out, 5.0f), si = powf(1.0f - n_in, 5.0f);
      const float sh = powf(1.0f - h_out, 5.0f);
      const float base = (1.0f + (fd90 - 1.0f) * so) * (1.0f + (fd90 - 1.0f) * si);
      const float fss90 = h_out * h_out * s.rough;
      const float fss = (1.0f + (fss90 - 1.0f) * so) * (1.0f + (fss90 - 1.0f) * si);
      const float ss = 1.25f * (fss * (1.0f / (n_out + n_in) - 0.5f) + 0.5f);
      const float kdiff = (1.0f - s.st) * (1.0f - s.metallic) *
                          ((base * (1.0f - s.subsurface) + ss * s.subsurface) / 3.14159265358979f) *
                          n_out;
      r = add(r, scale(s.kd, kdiff));
      const float3 ct = tint_of(s.kd);
      const float3 csheen = add(v3(1.0f - s.sheen_tint, 1.0f - s.sheen_tint, 1.0f - s.sheen_tint),
                                scale(ct, s.sheen_tint));
      r = add(r, scale(csheen, (1.0f - s.metallic) * s.sheen * sh * n_out));
      if (n_in > 0.0f && h_out > 0.0f && n_h > 0.0f) {
        const float f0 = (s.eta - 1.0f) * (s.eta - 1.0f) / ((s.eta + 1.0f) * (s.eta + 1.0f));
        const float3 sc = add(v3(1.0f - s.spec_tint, 1.0f - s.spec_tint, 1.0f - s.spec_tint),
                              scale(ct, s.spec_tint));
        const float3 c0 = add(scale(sc, s.specular * f0 * (1.0f - s.metallic)), scale(s.kd, s.metallic));
        const float sw = 1.0f - s.st * (1.0f - s.metallic);
        const float fw = schlick5(h_out);
        const float dg = gtr2(to_local(h, n), d.ax, d.ay) * smith_gtr2(to_local(v, n), d.ax, d.ay) *
                         smith_gtr2(to_local(l, n), d.ax, d.ay);
        const float k = sw * dg / (4.0f * n_in);
        r = add(r, v3((c0.x + (1.0f - c0.x) * fw) * k, (c0.y + (1.0f - c0.y) * fw) * k,
                      (c0.z + (1.0f - c0.z) * fw) * k));
      }
      if (n_in > 0.0f && n_h > 0.0f) {
        const float fc = 0.04f + 0.96f * schlick5(h_out);
        const float gc = smith_gtr2(to_local(v, n), 0.25f, 0.25f) *
                         smith_gtr2(to_local(l, n), 0.25f, 0.25f);
        const float c = d.cc * fc * gtr1(n_h, d.ac) * gc / (4.0f * n_in);
        r = add(r, v3(c, c, c));
      }
    }
    const float eta = dot(s.ng, v) > 0.0f ? s.eta : 1.0f / s.eta;
    const float fg = fresnel_dielectric(h_in, eta);
    const float dg = gtr2(to_local(h, n), d.ax, d.ay) * smith_gtr2(to_local(v, n), d.ax, d.ay) *
                     smith_gtr2(to_local(l, n), d.ax, d.ay);
    const float g = (1.0f - s.metallic) * s.st * (fg * dg) / (4.0f * fabsf(n_in));
    return add(r, scale(s.kd, g));
  }
  const float eta = dot(s.ng, v) > 0.0f ? s.eta : 1.0f / s.eta;
  float3 h = normalize(add(v, scale(l, eta)));
  if (dot(h, n) < 0.0f) h = scale(h, -1.0f);
  const float ef = 1.0f / (eta * eta);
  const float h_in = dot(h, v), h_out = dot(h, l), sd = h_in + eta * h_out;
  const float fg = fresnel_dielectric(h_in, eta);
  const float dg = gtr2(to_local(h, n), d.ax, d.ay) * smith_gtr2(to_local(v, n), d.ax, d.ay) *
                   smith_gtr2(to_local(l, n), d.ax, d.ay);
  const float t = (1.0f - s.metallic) * s.st *
                  (ef * (1.0f - fg) * dg * eta * eta * fabsf(h_out * h_in)) /
                  (fabsf(dot(n, v)) * sd * sd);
  return v3(sqrtf(s.kd.x) * t, sqrtf(s.kd.y) * t, sqrtf(s.kd.z) * t);
}

__device__ float ds_pdf(const Surface &s, float3 l, float3 v) {
  const DisneyLobes d = disney_lobes(s, v);
  const float3 n = d.n;
  const bool reflect = dot(s.ng, l) * dot(s.ng, v) > 0.0f;
  const float tot = d.dw + d.mw + d.gw + d.cw;
  float dp = d.dw / tot, mp = d.mw / tot, gp = d.gw / tot, cp = d.cw / tot;
  if (dot(s.ng, v) < 0.0f) {
    dp = mp = cp = 0.0f;
    if (gp > 0.0f) gp = 1.0f;
  }
  const float eta = dot(s.ng, v) > 0.0f ? s.eta : 1.0f / s.eta;
  if (reflect) {
    float3 h = normalize(add(v, l));
    if (dot(h, n) < 0.0f) h = scale(h, -1.0f);
    const float n_in = dot(n, v), n_h = dot(h, s.ns), h_in = dot(h, v), h_out = dot(h, l);
    if (dot(s.ng, l) >= 0.0f && dot(s.ng, v) >= 0.0f) {
      dp *= fmaxf(dot(n, v), 0.0f) / 3.14159265358979f;
      if (n_in > 0.0f)
        mp *= gtr2(to_local(h, n), d.ax, d.ay) * smith_gtr2(to_local(v, n), d.ax, d.ay) /
              (4.0f * n_in);
      else
        mp = 0.0f;
      if (n_h > 0.0f && h_out > 0.0f)
        cp *= gtr1(n_h, d.ac) * n_h / (4.0f * h_out);
      else
        cp = 0.0f;
    }
    gp *= fresnel_dielectric(h_in, eta) * gtr2(to_local(h, n), d.ax, d.ay) *
          smith_gtr2(to_local(v, n), d.ax, d.ay) / (4.0f * fabsf(n_in));
  } else {
    float3 h = normalize(add(v, scale(l, eta)));
    if (dot(h, n) < 0.0f) h = scale(h, -1.0f);
    const float h_in = dot(h, v), h_out = dot(h, l), sd = h_in + eta * h_out;
    const float dh = eta * eta * h_out / (sd * sd);
    gp *= (1.0f - fresnel_dielectric(h_in, eta)) * gtr2(to_local(h, n), d.ax, d.ay) *
          smith_gtr2(to_local(v, n), d.ax, d.ay) * fabsf(dh * h_in / dot(n, v));
  }
  return dp + mp + gp + cp;
}

// Heitz's visible-normal sample for an incoming direction below the
// surface flips it (shaderUtils.h:160-213).
__device__ inline float3 visible_normal(float3 li, float ax, float ay, float u1, float u2) {
  bool below = li.z < 0.0f;
  if (below) li = scale(li, -1.0f);
  const float3 hemi = normalize(v3(ax * li.x, ay * li.y, li.z));
  const float r = sqrtf(u1), phi = 2.0f * 3.14159265358979f * u2;
  const float t1 = r * cosf(phi);
  float t2 = r * sinf(phi);
  const float sm = (1.0f + hemi.z) * 0.5f;
  t2 = (1.0f - sm) * sqrtf(1.0f - t1 * t1) + sm * t2;
  const float3 disk = v3(t1, t2, sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2)));
  const float3 hn = to_world(disk, hemi);
  const float3 m = normalize(v3(ax * hn.x, ay * hn.y, fmaxf(0.0f, hn.z)));
  return below ? scale(m, -1.0f) : m;
}

__device__ inline float3 reflect_about(float3 v, float3 h) {
  return normalize(add(scale(v, -1.0f), scale(h, 2.0f * dot(v, h))));
}

// Glass lobe: reflect or refract about a visible normal, by the Fresnel
// term against u.  Total internal reflection returns `tir`.
__device__ inline float3 ds_glass(const Surface &s, const DisneyLobes &d, float3 v, float u,
                                  float ru, float rv, float3 tir) {
  float3 h = to_world(visible_normal(to_local(v, s.ns), d.ax, d.ay, ru, rv), s.ns);
  if (dot(h, d.n) < 0.0f) h = scale(h, -1.0f);
  const float h_in = dot(h, v);
  const float eta = dot(s.ng, v) > 0.0f ? s.eta : 1.0f / s.eta;
  if (u <= fresnel_dielectric(h_in, eta)) return reflect_about(v, h);
  const float hos = 1.0f - (1.0f - h_in * h_in) / (eta * eta);
  if (hos <= 0.0f) return tir;
  if (h_in < 0.0f) h = scale(h, -1.0f);
  const float ho = sqrtf(hos);
  return add(scale(v, -1.0f / eta), scale(h, fabsf(h_in) / eta - ho));
}

__device__ float3 ds_sample(const Surface &s, float3 v, float rw, float ru, float rv) {
  const DisneyLobes d = disney_lobes(s, v);
  if (dot(s.ng, v) >= 0.0f) {
    const float tot = d.dw + d.mw + d.gw + d.cw;
    const float dp = d.dw / tot, mp = d.mw / tot, gp = d.gw / tot;
    if (rw <= dp) {
      const float phi = 2.0f * 3.14159265358979f * ru;
      const float tmp = sqrtf(fminf(fmaxf(1.0f - rv, 0.0f), 1.0f));
      return to_world(v3(cosf(phi) * tmp, sinf(phi) * tmp, sqrtf(fminf(fmaxf(rv, 0.0f), 1.0f))), d.n);
    }
    if (rw <= dp + mp)
      return reflect_about(v, to_world(visible_normal(to_local(v, s.ns), d.ax, d.ay, ru, rv), s.ns));
    if (rw <= dp + mp + gp) {
      if (gp <= 0.0f) return v3(-2.0f, -2.0f, -2.0f);
      return ds_glass(s, d, v, (rw - (dp + mp)) / gp, ru, rv, v3(-2.0f, -2.0f, -2.0f));
    }
    const float a2 = d.ac * d.ac;
    const float ch = sqrtf(fmaxf(0.0f, (1.0f - powf(a2, 1.0f - ru)) / (1.0f - a2)));
    const float sh = sqrtf(fmaxf(1.0f - ch * ch, 0.0f)), az = 2.0f * 3.14159265358979f * rv;
    return reflect_about(v, to_world(v3(sh * cosf(az), sh * sinf(az), ch), d.n));
  }
  // From inside: the glass lobes only; total internal reflection gives the
  // reference's zero vector here.
  return ds_glass(s, d, v, rw, ru, rv, v3(0.0f, 0.0f, 0.0f));
}

// DS: the kernel instance carries the DisneyBSDF code (a mesh with a
// DisneyBSDF material); instances without it keep the register budget of
// the simpler families.
template <bool DS>
__device__ inline float3 bsdf_eval(const Surface &s, float3 l, float3 v) {
  if (DS && s.type == 1) return ds_eval(s, l, v);
  return s.type == 0 ? rp_eval(s, l, v) : s.type == 4 ? sg_eval(s, l, v) : ct_eval(s, l, v);
}
template <bool DS>
__device__ inline float bsdf_pdf(const Surface &s, float3 l, float3 v) {
  if (DS && s.type == 1) return ds_pdf(s, l, v);
  return s.type == 0 ? rp_pdf(s, l, v) : ct_pdf(s, l, v);
}
// The material's sampler from the stream at counter c: roughPlastic draws
// two uniforms, the others three.  x = -2 marks no sample.
template <bool DS>
__device__ inline float3 bsdf_sample(const Surface &s, float3 v, uint32_t seed, uint32_t key,
                                     uint32_t c) {
  const float r0 = rng_uniform(seed, key, c), r1 = rng_uniform(seed, key, c + 1);
  if (s.type == 0) return rp_sample(s, v, r0, r1);
  const float r2 = rng_uniform(seed, key, c + 2);
  if (DS && s.type == 1) return ds_sample(s, v, r0, r1, r2);
  return ct_sample(s, v, r0, r1, r2);
}

// Cube face and (u, v) of a direction (EnvDirecToUV, later tests win).
__device__ inline int dir_to_face(float3 d, float &u, float &v) {
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  float m = ax, uc = d.z, vc = d.y;
  int f = 0;
  if (d.x > 0.0f && ax >= ay && ax >= az) { m = ax; uc = d.z; vc = d.y; f = 0; }
  if (!(d.x > 0.0f) && ax >= ay && ax >= az) { m = ax; uc = -d.z; vc = d.y; f = 1; }
  if (d.y > 0.0f && ay >= ax && ay >= az) { m = ay; uc = d.x; vc = d.z; f = 2; }
  if (!(d.y > 0.0f) && ay >= ax && ay >= az) { m = ay; uc = d.x; vc = -d.z; f = 3; }
  if (!(d.z > 0.0f) && az >= ax && az >= ay) { m = az; uc = d.x; vc = d.y; f = 4; }
  if (d.z > 0.0f && az >= ax && az >= ay) { m = az; uc = -d.x; vc = d.y; f = 5; }
  u = 0.5f * (uc / m + 1.0f);
  v = 0.5f * (vc / m + 1.0f);
  return f;
}

// EnvUVToDirec (cubeMap.cu:94-110).
__device__ inline float3 face_to_dir(int f, float u, float v) {
  const float uc = 2.0f * u - 1.0f, vc = 2.0f * v - 1.0f;
  float3 d;
  switch (f) {
    case 0: d = v3(1.0f, vc, uc); break;
    case 1: d = v3(-1.0f, vc, -uc); break;
    case 2: d = v3(uc, 1.0f, vc); break;
    case 3: d = v3(uc, -1.0f, -vc); break;
    case 4: d = v3(uc, vc, -1.0f); break;
    default: d = v3(-uc, vc, 1.0f); break;
  }
  return normalize(d);
}

// Table layout (env_table_doubles): pmf[6] cdf[7] face_lum[6] | per face.
__host__ __device__ inline size_t env_face_doubles(int n) {
  return (size_t)n * (n + 1) + (size_t)n * n + (n + 1) + n;
}
constexpr int kTabHead = 19;

// First index i in [0, count) with cdf[i] > r (count when none), as the
// linear searches of cubeMap.cu:115-164 find it.
__device__ inline int first_above(const double *cdf, int count, float r) {
  int lo = 0, hi = count;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cdf[mid] > (double)r) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// __direct_callable__sampleEnvironmapLight (cubeMap.cu:166-174).
__device__ float3 env_sample(const double *tab, int n, float r0, float r1, float r2) {
#pragma clang fp contract(off)
  const int f = min(max(first_above(tab + 6, 6, r0) - 1, 0), 5);
  const double *base = tab + kTabHead + (size_t)f * env_face_doubles(n);
  const double *cdf_rows = base, *cdf_m = base + (size_t)n * (n + 1) + (size_t)n * n;
  const int y = min(max(first_above(cdf_m, n, r2) - 1, 0), n - 1);
  double dy = (double)r2 - cdf_m[y];
  if (cdf_m[y + 1] - cdf_m[y] > 0) dy /= (cdf_m[y + 1] - cdf_m[y]);
  const double *cdf = cdf_rows + (size_t)y * (n + 1);
  const int x = min(max(first_above(cdf, n, r1) - 1, 0), n - 1);
  double dx = (double)r1 - cdf[x];
  if (cdf[x + 1] - cdf[x] > 0) dx /= (cdf[x + 1] - cdf[x]);
  return face_to_dir(f, (float)((x + dx) / n), (float)((y + dy) / n));
}

// __direct_callable__pdf (cubeMap.cu:178-186) with pdf_envmap
// (cubeMap.cu:9-19), in the same FP32 operations: a density per unit face
// (u, v) area times the face probability.
__device__ float env_pdf(const double *tab, int n, float3 d) {
#pragma clang fp contract(off)
  float u, v;
  const int f = dir_to_face(d, u, v);
  const double *base = tab + kTabHead + (size_t)f * env_face_doubles(n);
  const double *pdf_rows = base + (size_t)n * (n + 1);
  const double *pdf_m = pdf_rows + (size_t)n * n + (n + 1);
  const int x = (int)fminf(fmaxf(u * (float)n, 0.0f), (float)(n - 1));
  const int y = (int)fminf(fmaxf(v * (float)n, 0.0f), (float)(n - 1));
  const float py = (float)pdf_m[y], px = (float)pdf_rows[(size_t)y * n + x];
  const float p = py * px * (float)n * (float)n;
  return (float)tab[f] * p;
}

__device__ inline bool finite3(float3 c) {
  return !(isnan(c.x) || isnan(c.y) || isnan(c.z));
}

// Direct light at a first hit p seen along d (unit), averaged over the
// pixel's samples.
template <bool DS>
__device__ float3 direct_light(const CastArgs &a, const Surface &s, float3 p, float3 d,
                               uint32_t pixel, int *stack) {
  const float3 v = scale(d, -1.0f);
  float acc[3] = {0.0f, 0.0f, 0.0f};
  for (int k = 0; k < a.direct_samples; ++k) {
    const uint32_t c = (uint32_t)k * 8u;
    // Environment-importance sample, power-heuristic weight.
    {
      const float3 l = env_sample(a.env_tab, a.light_n, rng_uniform(a.seed, pixel, c),
                                  rng_uniform(a.seed, pixel, c + 1), rng_uniform(a.seed, pixel, c + 2));
      const float pe = env_pdf(a.env_tab, a.light_n, l);
      if (pe > 1e-4f && !occluded(a.nodes, a.tris, p, l, stack)) {
        const float pb = bsdf_pdf<DS>(s, l, v);
        const float3 f = bsdf_eval<DS>(s, l, v);
        const float3 le = env_radiance(a.light_faces, a.light_n, l);
        const float w = pe / fmaxf(pe * pe + pb * pb, 1e-8f);
        const float3 r = v3(f.x * le.x * w, f.y * le.y * w, f.z * le.z * w);
        if (finite3(r)) {
          acc[0] += r.x;
          acc[1] += r.y;
          acc[2] += r.z;
        }
      }
    }
    // BSDF sample, power-heuristic weight.
    {
      const float3 l = bsdf_sample<DS>(s, v, a.seed, pixel, c + 3);
      if (l.x != -2.0f) {
        const float pb = bsdf_pdf<DS>(s, l, v);
        const float pe = env_pdf(a.env_tab, a.light_n, l);
        if (pe > 1e-4f && !occluded(a.nodes, a.tris, p, l, stack)) {
          const float3 f = bsdf_eval<DS>(s, l, v);
          const float3 le = env_radiance(a.light_faces, a.light_n, l);
          const float w = pb / fmaxf(pe * pe + pb * pb, 1e-8f);
          const float3 r = v3(f.x * le.x * w, f.y * le.y * w, f.z * le.z * w);
          if (finite3(r)) {
            acc[0] += r.x;
            acc[1] += r.y;
            acc[2] += r.z;
          }
        }
      }
    }
  }
  const float inv = 1.0f / (float)a.direct_samples;
  return v3(acc[0] * inv, acc[1] * inv, acc[2] * inv);
}

// Sampling tables (Environment_Map::init_sampling_dist,
// EnvironmentMap.cpp:153-178, and make_table_dist_2d, table_dist.cpp:43-118).
// Table row r, column c of a face holds texel (r, c), weighted by
// TexelCoordSolidAngle evaluated at ((r + 0.5) / n, (c + 0.5) / n) as the
// reference calls it (EnvironmentMap.cpp:128-147, 167-174).
// Pass 1: one workgroup per (face, row), one lane per texel; the row's
// prefix sum runs in order on lane 0.  The row integral and the row's raw
// luminance sum are parked in the face's marginal arrays for pass 2.
__global__ __launch_bounds__(128) void env_rows_kernel(const float *__restrict__ faces, int n,
                                                      double *__restrict__ tab) {
#pragma clang fp contract(off)
  const int f = blockIdx.x / n, r = blockIdx.x % n, c = threadIdx.x;
  double *base = tab + kTabHead + (size_t)f * env_face_doubles(n);
  double *cr = base + (size_t)r * (n + 1);
  double *pr = base + (size_t)n * (n + 1) + (size_t)r * n;
  double *cdf_m = base + (size_t)n * (n + 1) + (size_t)n * n, *pdf_m = cdf_m + (n + 1);
  __shared__ double fv[128], lv[128];
  if (c < n) {
    const double U = (2.0 * ((double)((r + 0.5) / n) + 0.5) / (double)n) - 1.0;
    const double V = (2.0 * ((double)((c + 0.5) / n) + 0.5) / (double)n) - 1.0;
    const double inv = 1.0 / n;
    auto area = [](double x, double y) { return atan2(x * y, sqrt(x * x + y * y + 1.0)); };
    const double sa = area(U - inv, V - inv) - area(U - inv, V + inv) - area(U + inv, V - inv) +
                      area(U + inv, V + inv);
    const float saf = (float)sa;
    const float kScale = 2.0f * 3.14159265358979323846f / 3.0f;
    const float *px = faces + (((size_t)f * n + r) * n + c) * 3;
    const float3 t = v3(px[0], px[1], px[2]);
    const float3 q = v3(fabsf(t.x) * saf / kScale, fabsf(t.y) * saf / kScale,
                        fabsf(t.z) * saf / kScale);
    fv[c] = (double)luminance(q);
    lv[c] = (double)luminance(t);
  }
  __syncthreads();
  if (c == 0) {
    double acc = 0.0, lum = 0.0;
    cr[0] = 0.0;
    for (int x = 0; x < n; ++x) {
      acc += fv[x];
      cr[x + 1] = acc;
      lum += lv[x];
    }
    pdf_m[r] = acc;  // row integral, read by pass 2
    cdf_m[r] = lum;  // row luminance, read by pass 2
  }
  __syncthreads();
  if (c < n) {
    const double integral = cr[n];
    if (integral > 0) {
      cr[c] /= integral;
      pr[c] = fv[c] / integral;
    } else {
      cr[c] = 0.0;
      pr[c] = 0.0;
    }
  }
}

// Pass 2: one workgroup per face: marginals in row order, the face's
// luminance, and each row's last CDF entry set to 1 (0 for a black row).
__global__ __launch_bounds__(128) void env_marginals_kernel(int n, double *__restrict__ tab) {
#pragma clang fp contract(off)
  const int f = blockIdx.x, r = threadIdx.x;
  double *base = tab + kTabHead + (size_t)f * env_face_doubles(n);
  double *cdf_m = base + (size_t)n * (n + 1) + (size_t)n * n, *pdf_m = cdf_m + (n + 1);
  __shared__ double row_int[128], row_lum[128];
  if (r < n) {
    row_int[r] = pdf_m[r];
    row_lum[r] = cdf_m[r];
  }
  __syncthreads();
  if (r == 0) {
    double lum = 0.0, acc = 0.0;
    cdf_m[0] = 0.0;
    for (int y = 0; y < n; ++y) {
      acc += row_int[y];
      cdf_m[y + 1] = acc;
      lum += row_lum[y];
    }
    const double total = acc;
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
    tab[13 + f] = lum;
  }
  if (r < n) {
    double *cr = base + (size_t)r * (n + 1);
    cr[n] = cr[n - 1] > 0 ? 1.0 : 0.0;
  }
}

// Face distribution (EnvironmentMap.cpp:180-189, make_table_dist_1d,
// table_dist.cpp:4-27); an all-black map gets the uniform distribution.
__global__ void env_face_table_kernel(double *__restrict__ tab) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  double total = 0.0, f[6];
  for (int i = 0; i < 6; ++i) total += tab[13 + i];
  for (int i = 0; i < 6; ++i) f[i] = total > 0 ? tab[13 + i] / total : 0.0;
  double cdf[7];
  cdf[0] = 0.0;
  for (int i = 0; i < 6; ++i) cdf[i + 1] = cdf[i] + f[i];
  const double t = cdf[6];
  for (int i = 0; i < 6; ++i) {
    tab[i] = t > 0 ? f[i] / t : 1.0 / 6.0;
    tab[6 + i] = t > 0 ? cdf[i] / t : i / 6.0;
  }
  tab[12] = t > 0 ? t : 1.0;
}

// One 8 x 8 pixel tile per 64-lane workgroup (coherent primary rays).
// DIRECT: the direct-light estimate is computed and stored (a separate
// instance, so the plain cast keeps its lower register count).
// At most 128 VGPRs (4 waves per SIMD, the LDS stack allows 5).
template <bool DIRECT, bool DS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void gbuffer_cast_kernel(
    CastArgs a) {
  __shared__ int stack_lds[kBvhMaxDepth * 64];
  const int lane = threadIdx.x;
  const int tiles_x = (a.width + 7) >> 3;
  const int ix = (int)(blockIdx.x % (unsigned)tiles_x) * 8 + (lane & 7);
  const int iy = (int)(blockIdx.x / (unsigned)tiles_x) * 8 + (lane >> 3);
  if (ix >= a.width || iy >= a.height) return;
  int *stack = stack_lds + lane;

  const float sx = ((float)ix - a.cx) / a.fx, sy = ((float)iy - a.cy) / a.fy;
  const float3 cd = v3(a.cam_dir), ch = v3(a.cam_h), cv = v3(a.cam_v);
  float3 d = normalize(add(add(cd, scale(ch, sx)), scale(cv, sy)));
  float3 o = v3(a.cam_pos);

  // What the pixel stores; defaults are the cleared state of a path that
  // ended on a mirror without a first hit.
  float3 pos = v3(0.0f, 0.0f, 0.0f), nrm = v3(0.0f, 0.0f, 0.0f), refl = nrm;
  float3 kd = nrm, ks = nrm, direct = nrm;
  float rough = 0.0f;
  unsigned char flag = 0;
  constexpr bool want_direct = DIRECT;
  for (int depth = 1;; ++depth) {
    const Hit h = trace(a.nodes, a.tris, o, d, stack);
    if (h.prim == INT_MAX) {
      nrm = scale(d, -1.0f);
      refl = nrm;
      pos = v3(-100000.0f, -100000.0f, -100000.0f);
      if (depth == 1 && a.env_faces) kd = env_radiance(a.env_faces, a.env_n, d);
      // The first ray's miss colour is also its direct term (camera.cu:145).
      direct = kd;
      ks = v3(0.0f, 0.0f, 0.0f);
      rough = 1.0f;
      break;
    }
    const int mid = a.tri_material ? a.tri_material[h.prim] : 0;
    MaterialRec mat = a.materials[mid];
    const float4 *rec = reinterpret_cast<const float4 *>(a.tris + h.slot);
    const float3 ng = normalize(cross(v3(rec[1]), v3(rec[2])));
    float3 ns = ng;
    if (a.normals) {
      const int *tri = a.indices + 3 * (size_t)h.prim;
      const float w = 1.0f - h.u - h.v;
      const float3 n0 = v3(a.normals + 3 * (size_t)tri[0]);
      const float3 n1 = v3(a.normals + 3 * (size_t)tri[1]);
      const float3 n2 = v3(a.normals + 3 * (size_t)tri[2]);
      ns = normalize(add(add(scale(n0, w), scale(n1, h.u)), scale(n2, h.v)));
    }
    const float3 p = add(add(o, scale(d, h.t)), v3(0.001f, 0.001f, 0.001f));
    if (mat.type == NWPRT_MATERIAL_MIRROR) {
      if (depth >= a.max_bounce) break;
      o = p;
      d = normalize(reflect(d, ns));
      continue;
    }
    if (a.mat_tex)
      mat = textured(mat, a.mat_tex + (size_t)mid * kTexFields,
                     hit_texcoord(a.texcoords, a.indices, h.prim, h.u, h.v));
    pos = p;
    nrm = ns;
    refl = normalize(reflect(d, ns));
    kd = v3(mat.kd);
    ks = v3(mat.ks);
    if (mat.type == 1) {  // DisneyBSDF's PACKINFO (disneyBSDF.cu:10-25)
      const float3 ct = tint_of(kd);
      const float f0 = (mat.eta - 1.0f) * (mat.eta - 1.0f) / ((mat.eta + 1.0f) * (mat.eta + 1.0f));
      const float3 sc = add(v3(1.0f - mat.specular_tint, 1.0f - mat.specular_tint,
                               1.0f - mat.specular_tint),
                            scale(ct, mat.specular_tint));
      ks = add(scale(sc, mat.specular * f0 * (1.0f - mat.metallic)), scale(kd, mat.metallic));
    }
    rough = mat.roughness;
    flag = 1;
    if constexpr (want_direct) {
      const Surface sf = make_surface(mat, ns, ng);
      direct = direct_light<DS>(a, sf, p, d, (uint32_t)(iy * a.width + ix), stack);
    }
    break;
  }

  const size_t px = (size_t)iy * a.width + ix;
  const nwprt_gbuffer &g = a.gb;
  if (g.position)
    reinterpret_cast<float4 *>(g.position)[px] =
        make_float4((pos.x - a.lower[0]) / a.box[0], (pos.y - a.lower[1]) / a.box[1],
                    (pos.z - a.lower[2]) / a.box[2], 1.0f);
  if (g.normal) reinterpret_cast<float4 *>(g.normal)[px] = enc(nrm);
  if (g.reflect) reinterpret_cast<float4 *>(g.reflect)[px] = enc(refl);
  if (g.kd) reinterpret_cast<float4 *>(g.kd)[px] = make_float4(kd.x, kd.y, kd.z, 1.0f);
  if (g.ks) reinterpret_cast<float4 *>(g.ks)[px] = make_float4(ks.x, ks.y, ks.z, 1.0f);
  if (g.roughness) g.roughness[px] = rough;
  if (g.env_flag) g.env_flag[px] = flag;
  if constexpr (want_direct)
    reinterpret_cast<float4 *>(g.direct)[px] = make_float4(direct.x, direct.y, direct.z, 1.0f);
}

// ------------------------------------------------------------------ SH-PRT
// The reference's SH-PRT render (RenderEngine::shRender,
// RenderEngine.cpp:459-486, and __closesthit__SHradiance,
// devicePrograms.cu:312-501): each hit triangle's three vertex colours are
// |sum_k R_k (x) B_k| with R = T_v L the vertex's light response (transfer
// matrix times the environment's SH vector) and B the Monte Carlo SH
// projection of the BSDF at the vertex (normal and view direction of the
// vertex, f / pdf weights); the pixel interpolates them barycentrically.
// Here the vertex colours are computed once per distinct (vertex, material)
// pair of the triangle corners instead of once per pixel: the same
// expectation.

constexpr int kShCoeffs = (kShMaxDegree + 1) * (kShMaxDegree + 1);

// Real spherical harmonics as pbrt defines them (shaderUtils.h:230-337):
// index l (l + 1) + m; associated Legendre functions with the
// Condon-Shortley phase; K = sqrt((2l + 1) / (4 pi) (l - |m|)! / (l + |m|)!);
// sqrt(2) K P_l^|m| sin(|m| phi) for m < 0, K P_l^0, sqrt(2) K P_l^m
// cos(m phi) for m > 0.  Degree 8 (81 values), fully unrolled.
__device__ inline void sh_eval(float3 w, float (&y)[kShCoeffs]) {
  constexpr int L = kShMaxDegree;
  auto idx = [](int l, int m) { return l * l + l + m; };
  const float x = w.z;
#pragma unroll
  for (int k = 0; k < kShCoeffs; ++k) y[k] = 0.0f;
  y[idx(0, 0)] = 1.0f;
  y[idx(1, 0)] = x;
#pragma unroll
  for (int l = 2; l <= L; ++l)
    y[idx(l, 0)] = ((2 * l - 1) * x * y[idx(l - 1, 0)] - (l - 1) * y[idx(l - 2, 0)]) / l;
  {
    float neg = -1.0f, dfact = 1.0f;
    const float xroot = sqrtf(fmaxf(0.0f, 1.0f - x * x));
    float xpow = xroot;
#pragma unroll
    for (int l = 1; l <= L; ++l) {
      y[idx(l, l)] = neg * dfact * xpow;
      neg = -neg;
      dfact *= 2 * l + 1;
      xpow *= xroot;
    }
  }
#pragma unroll
  for (int l = 2; l <= L; ++l) y[idx(l, l - 1)] = x * (2 * l - 1) * y[idx(l - 1, l - 1)];
#pragma unroll
  for (int l = 3; l <= L; ++l)
#pragma unroll
    for (int m = 1; m <= l - 2; ++m)
      y[idx(l, m)] = ((2 * (l - 1) + 1) * x * y[idx(l - 1, m)] - (l - 1 + m) * y[idx(l - 2, m)]) /
                     (l - m);
  float sins[L + 1], coss[L + 1];
  const float xy = sqrtf(fmaxf(0.0f, 1.0f - x * x));
  {
    const float sp = xy == 0.0f ? 0.0f : w.y / xy, cp = xy == 0.0f ? 1.0f : w.x / xy;
    float si = 0.0f, ci = 1.0f;
#pragma unroll
    for (int i = 0; i <= L; ++i) {
      sins[i] = xy == 0.0f ? 0.0f : si;
      coss[i] = xy == 0.0f ? 1.0f : ci;
      const float os = si;
      si = si * cp + ci * sp;
      ci = ci * cp - os * sp;
    }
  }
  const float sqrt2 = 1.41421356237f;
#pragma unroll
  for (int l = 0; l <= L; ++l) {
#pragma unroll
    for (int m = -l; m <= l; ++m) {
      const int am = m < 0 ? -m : m;
      float inv = 1.0f;  // (l - |m|)! / (l + |m|)!
#pragma unroll
      for (int t = l - am + 1; t <= l + am; ++t) inv /= (float)t;
      const float k = sqrtf((2.0f * l + 1.0f) * 0.0795774715459f * inv);
      if (m < 0) y[idx(l, m)] = sqrt2 * k * y[idx(l, am)] * sins[am];
    }
    y[idx(l, 0)] *= sqrtf((2.0f * l + 1.0f) * 0.0795774715459f);
#pragma unroll
    for (int m = 1; m <= l; ++m) {
      float inv = 1.0f;
#pragma unroll
      for (int t = l - m + 1; t <= l + m; ++t) inv /= (float)t;
      y[idx(l, m)] *= sqrt2 * sqrtf((2.0f * l + 1.0f) * 0.0795774715459f * inv) * coss[m];
    }
  }
}

// R[v][k] = sum_j T[v][k][j] (x) L[j].  One wavefront per vertex walks its
// nc rows; a row's nc x 3 floats are read by the 64 lanes together
// (coalesced) and multiplied by L flattened the same way, then reduced
// across the wave per channel.  Four vertices per 256-thread workgroup.
__device__ inline float wave_sum(float x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// A row of the transfer matrix is nc x 3 <= 243 floats: at most 4 per lane.
// The loop keeps kShRowsInFlight rows' loads issued ahead of the row being
// reduced, so that a wave has several rows of HBM reads outstanding while
// its cross-lane sums run.
constexpr int kShRowSlots = 4;
constexpr int kShRowsInFlight = 4;
static_assert((kShMaxDegree + 1) * (kShMaxDegree + 1) * 3 <= 64 * kShRowSlots,
              "a transfer-matrix row must fit the wave's load slots");

__global__ __launch_bounds__(256) void sh_light_response_kernel(ShArgs a) {
  const int nc = (a.degree + 1) * (a.degree + 1), row = nc * 3;
  const int lane = threadIdx.x & 63;
  const long long v = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (v >= a.n_vertices) return;  // whole waves leave together; no barriers below
  const float *t = a.transfer + (size_t)v * nc * row;
  float *o = a.light_response + (size_t)v * row;
  float l[kShRowSlots];
  int c[kShRowSlots];
#pragma unroll
  for (int s = 0; s < kShRowSlots; ++s) {
    const int f = lane + 64 * s;
    l[s] = f < row ? a.light_sh[f] : 0.0f;
    c[s] = f % 3;
  }
  float buf[kShRowsInFlight][kShRowSlots];
  auto load_row = [&](int k, float(&r)[kShRowSlots]) {
#pragma unroll
    for (int s = 0; s < kShRowSlots; ++s) {
      const int f = lane + 64 * s;
      r[s] = (k < nc && f < row) ? t[(size_t)k * row + f] : 0.0f;
    }
  };
#pragma unroll
  for (int d = 0; d < kShRowsInFlight; ++d) load_row(d, buf[d]);
  for (int k0 = 0; k0 < nc; k0 += kShRowsInFlight) {
#pragma unroll
    for (int d = 0; d < kShRowsInFlight; ++d) {
      float cur[kShRowSlots];
#pragma unroll
      for (int s = 0; s < kShRowSlots; ++s) cur[s] = buf[d][s];
      load_row(k0 + kShRowsInFlight + d, buf[d]);
      const int k = k0 + d;
      if (k >= nc) continue;  // wave-uniform
      float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f;
#pragma unroll
      for (int s = 0; s < kShRowSlots; ++s) {
        const float x = cur[s] * l[s];
        acc0 += c[s] == 0 ? x : 0.0f;
        acc1 += c[s] == 1 ? x : 0.0f;
        acc2 += c[s] == 2 ? x : 0.0f;
      }
      const float r0 = wave_sum(acc0), r1 = wave_sum(acc1), r2 = wave_sum(acc2);
      if (lane == 0) {
        o[3 * k] = r0;
        o[3 * k + 1] = r1;
        o[3 * k + 2] = r2;
      }
    }
  }
}

// Colour of each distinct (vertex, material) pair of the triangle corners:
// the vertex's BSDF SH projection (normal and view direction of the vertex,
// devicePrograms.cu:446-478) dotted with its light response, then |.| per
// channel.  kShPairLanes lanes share one pair and take every kShPairLanes-th
// sample, so that the few tens of thousands of pairs of a mesh still fill the
// chip; their partial sums meet in a shuffle reduction.
constexpr int kShPairLanes = 16;

template <bool DS>
__global__ __launch_bounds__(64) void sh_corner_kernel(ShArgs a) {
  const int g = threadIdx.x & (kShPairLanes - 1);
  const long long u = (long long)blockIdx.x * (64 / kShPairLanes) + threadIdx.x / kShPairLanes;
  // Lanes past the last pair stay for the shuffles and sample nothing.
  const bool live = u < a.n_keys;
  const int vi = live ? a.key_vertex[u] : 0;
  const int mid = live ? a.key_material[u] : 0;
  const MaterialRec mat = a.materials[mid];
  const float3 pv = v3(a.vertices + 3 * (size_t)vi);
  const float3 n = normalize(v3(a.normals + 3 * (size_t)vi));
  const float3 view = normalize(sub(v3(a.cam_pos), pv));
  const Surface sf = make_surface(mat, n, n);
  const int nc = (a.degree + 1) * (a.degree + 1);
  const float *lr = a.light_response + (size_t)vi * nc * 3;
  const uint32_t seed = a.seed ^ (0x9E3779B9u * (uint32_t)(mid + 1));
  float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f;
  int valid = 0;
  for (int k = g; live && k < a.samples; k += kShPairLanes) {
    const uint32_t q = (uint32_t)k * 4u;
    const float3 l = bsdf_sample<DS>(sf, view, seed, (uint32_t)vi, q);
    if (l.x == -2.0f) continue;
    const float pdf = bsdf_pdf<DS>(sf, l, view);
    if (pdf < 1e-6f) continue;
    const float3 f = bsdf_eval<DS>(sf, l, view);
    if (!finite3(f)) continue;
    float y[kShCoeffs];
    sh_eval(l, y);
    float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f;
#pragma unroll
    for (int j = 0; j < kShCoeffs; ++j) {
      if (j < nc) {
        d0 = fmaf(y[j], lr[3 * j], d0);
        d1 = fmaf(y[j], lr[3 * j + 1], d1);
        d2 = fmaf(y[j], lr[3 * j + 2], d2);
      }
    }
    acc0 += f.x / pdf * d0;
    acc1 += f.y / pdf * d1;
    acc2 += f.z / pdf * d2;
    ++valid;
  }
#pragma unroll
  for (int o = kShPairLanes / 2; o > 0; o >>= 1) {
    acc0 += __shfl_xor(acc0, o, kShPairLanes);
    acc1 += __shfl_xor(acc1, o, kShPairLanes);
    acc2 += __shfl_xor(acc2, o, kShPairLanes);
    valid += __shfl_xor(valid, o, kShPairLanes);
  }
  if (!live || g != 0) return;
  const float inv = valid > 0 ? 1.0f / (float)valid : 0.0f;
  float *o = a.key_color + 3 * (size_t)u;
  o[0] = fabsf(acc0 * inv);
  o[1] = fabsf(acc1 * inv);
  o[2] = fabsf(acc2 * inv);
}

// Per pixel: the primary ray's first hit interpolates its triangle's corner
// colours; a miss shows the view environment (devicePrograms.cu:519-522).
__global__ __launch_bounds__(64) void sh_pixel_kernel(ShArgs a) {
  __shared__ int stack_lds[kBvhMaxDepth * 64];
  const int lane = threadIdx.x;
  const int tiles_x = (a.width + 7) >> 3;
  const int ix = (int)(blockIdx.x % (unsigned)tiles_x) * 8 + (lane & 7);
  const int iy = (int)(blockIdx.x / (unsigned)tiles_x) * 8 + (lane >> 3);
  if (ix >= a.width || iy >= a.height) return;
  const float sx = ((float)ix - a.cx) / a.fx, sy = ((float)iy - a.cy) / a.fy;
  const float3 d = normalize(add(add(v3(a.cam_dir), scale(v3(a.cam_h), sx)), scale(v3(a.cam_v), sy)));
  const Hit h = trace(a.nodes, a.tris, v3(a.cam_pos), d, stack_lds + lane);
  float3 c = v3(0.0f, 0.0f, 0.0f);
  if (h.prim == INT_MAX) {
    if (a.env_faces) c = env_radiance(a.env_faces, a.env_n, d);
  } else {
    const int32_t *ck = a.corner_key + 3 * (size_t)h.prim;
    const float *c0 = a.key_color + 3 * (size_t)ck[0], *c1 = a.key_color + 3 * (size_t)ck[1],
                *c2 = a.key_color + 3 * (size_t)ck[2];
    const float w = 1.0f - h.u - h.v;
    c = v3(w * c0[0] + h.u * c1[0] + h.v * c2[0], w * c0[1] + h.u * c1[1] + h.v * c2[1],
           w * c0[2] + h.u * c1[2] + h.v * c2[2]);
  }
  reinterpret_cast<float4 *>(a.color)[(size_t)iy * a.width + ix] = make_float4(c.x, c.y, c.z, 1.0f);
}


// ------------------------------------------------------------ SH transfer
// RenderEngine::shComputeTMatrix (RenderEngine.cpp:410-446) with its ray
// generation program __raygen__ComputeTMatrix (cameraShTmatrix.cu:87-178):
// per vertex, `samples` paths start at the vertex in uniform sphere
// directions; under each basis map i (at once the view, lighting and
// sampling environment) a path gathers one environment-importance light
// sample per non-mirror hit and the environment radiance where it leaves
// the scene; the radiance that came straight from the environment (or
// through one mirror) is removed and the rest is projected onto the
// filtered SH basis of the start direction.
//
// The path's control flow -- hits, BSDF continuations, Russian roulette --
// does not depend on the basis map (the light samples draw their own
// counters of the stream, and the throughput is map-independent), so it is
// traced once per (vertex, sample) and shared by all nc maps:
//   tm_paths_kernel    one lane per (vertex, sample): closest-hit rays,
//                      continuations and roulette; records each non-mirror
//                      hit (position, view direction, normals, material,
//                      throughput) and the exit ray;
//   tm_light_kernel    one lane per (vertex, sample, map), consecutive lanes
//                      on consecutive maps of one path: the light samples
//                      (shadow rays) at the recorded hits, the exit
//                      radiance, the removal of the direct part, giving the
//                      indirect radiance of the (sample, map) pair;
//   tm_project_kernel  one workgroup per vertex, one lane per (map, channel)
//                      column: T[k][i][c] = sum over samples, in sample
//                      order, of Y'_k(start direction) x indirect x 4 pi / N
//                      (Y' the filtered basis), with the reference's
//                      T += T / N for a NaN sample.
// Random counters per (vertex, sample): sample * 256 + {0, 1} for the start
// direction, + 16 (b + 1) + {0..2} for the light sample at trace b,
// {3..5} for the BSDF light sample (max_bounce 1), {6..8} for the
// continuation, 9 for the roulette.

constexpr uint32_t kTmStride = 256;
// Filtered SH basis of a start direction, padded to a 16-byte multiple.
constexpr int kTmRow = (kShCoeffs + 3) & ~3;

// A non-mirror hit of a path (64 bytes).  tag = material << 4 | trace index.
struct TmRecord {
  float p[3], v[3], ns[3], ng[3], lc[3];
  int tag;
};
// max_bounce 1 only, in the slot after the hit: the BSDF light sample
// (direction, pdf, BSDF value) and whether its shadow ray is clear.
struct TmBsdfLight {
  float l[3], pb, f[3];
  int visible;
  float pad[8];
};
static_assert(sizeof(TmRecord) == 64 && sizeof(TmBsdfLight) == 64, "record size");
// How a path ends (32 bytes): the exit ray's direction, throughput and BSDF
// pdf; info = records | traces << 4 | left the scene << 9 | mirror before
// any other hit << 10.
struct TmPath {
  float dir[3], lc[3], pdf;
  int info;
};

__host__ __device__ inline int tm_slots(int max_bounce) { return max_bounce == 1 ? 2 : max_bounce; }

__host__ __device__ inline size_t tm_align(size_t b) { return (b + 255) & ~(size_t)255; }

struct TmWork {
  TmRecord *rec;  // [pair][slots]
  TmPath *path;   // [pair]
  float *y;       // [pair][kTmRow]: filtered basis of the start direction
  float *ind;     // [pair][map][3]: indirect radiance
  float2 *tc;     // [pair][slots]: a record's texture coordinates (textured meshes)
};
__host__ __device__ inline TmWork tm_work(void *base, long long pairs, int slots, int nc) {
  char *p = (char *)base;
  TmWork w;
  w.rec = (TmRecord *)p;
  p += tm_align((size_t)pairs * slots * sizeof(TmRecord));
  w.path = (TmPath *)p;
  p += tm_align((size_t)pairs * sizeof(TmPath));
  w.y = (float *)p;
  p += tm_align((size_t)pairs * kTmRow * sizeof(float));
  w.ind = (float *)p;
  p += tm_align((size_t)pairs * nc * 3 * sizeof(float));
  w.tc = (float2 *)p;  // only addressed for textured meshes
  return w;
}

// sample_sphere (shaderUtils.h:151-158).
__device__ inline float3 tm_sphere(float z1, float z2) {
  const float theta = 2.0f * 3.14159265358979f * z1, phi = acosf(1.0f - 2.0f * z2);
  const float sp = sinf(phi);
  return v3(sp * cosf(theta), sp * sinf(theta), cosf(phi));
}

template <bool DS>
__device__ inline float3 tm_bsdf_sample(const Surface &s, float3 v, uint32_t seed, uint32_t vtx,
                                        uint32_t c) {
  return bsdf_sample<DS>(s, v, seed, vtx, c);
}

// The surface of a path vertex; tc: its texture coordinates (read only for
// textured meshes).
__device__ inline Surface tm_surface(const TmArgs &a, int mat, float3 ns, float3 ng, float2 tc) {
  if (a.mat_tex) return make_surface(textured(a.materials[mat], a.mat_tex + (size_t)mat * kTexFields, tc), ns, ng);
  return make_surface(a.materials[mat], ns, ng);
}

template <bool DS>
__global__ __launch_bounds__(64) void tm_paths_kernel(TmArgs a) {
  __shared__ int stack_lds[kBvhMaxDepth * 64];
  const int lane = threadIdx.x;
  const long long pair = (long long)blockIdx.x * 64 + lane;
  if (pair >= a.n_verts * a.samples) return;  // no barriers below
  int *stack = stack_lds + lane;
  const int slots = tm_slots(a.max_bounce);
  const TmWork w = tm_work(a.work, a.n_verts * a.samples, slots, (a.degree + 1) * (a.degree + 1));
  TmRecord *rec = w.rec + pair * slots;
  const long long vl = pair / a.samples;
  const int s = (int)(pair - vl * a.samples);
  const uint32_t vtx = (uint32_t)(a.v_begin + vl);
  const uint32_t base = (uint32_t)s * kTmStride;
  float3 d = tm_sphere(rng_uniform(a.seed, vtx, base), rng_uniform(a.seed, vtx, base + 1));
  float3 o = v3(a.vertices + 3 * (size_t)(a.v_begin + vl));
  float3 lc = v3(1.0f, 1.0f, 1.0f);
  float pdf = 0.0f;
  bool first_hit = false, mirror_first = false, left = false, done = false;
  int depth = 0, nrec = 0;
  TmPath out{};
  for (;;) {
    const uint32_t c = base + 16u * (uint32_t)(depth + 1);
    const float3 rd = normalize(d);
    const Hit h = trace(a.nodes, a.tris, o, d, stack);
    if (h.prim == INT_MAX) {
      out.dir[0] = d.x, out.dir[1] = d.y, out.dir[2] = d.z;
      out.lc[0] = lc.x, out.lc[1] = lc.y, out.lc[2] = lc.z;
      out.pdf = pdf;
      left = done = true;
    } else {
      const int mat = a.tri_material ? a.tri_material[h.prim] : 0;
      const float4 *tr = reinterpret_cast<const float4 *>(a.tris + h.slot);
      const float3 ng = normalize(cross(v3(tr[1]), v3(tr[2])));
      float3 ns = ng;
      if (a.normals) {
        const int *tri = a.indices + 3 * (size_t)h.prim;
        const float ww = 1.0f - h.u - h.v;
        ns = normalize(add(add(scale(v3(a.normals + 3 * (size_t)tri[0]), ww),
                               scale(v3(a.normals + 3 * (size_t)tri[1]), h.u)),
                           scale(v3(a.normals + 3 * (size_t)tri[2]), h.v)));
      }
      const float3 p = add(add(o, scale(rd, h.t)), v3(0.001f, 0.001f, 0.001f));
      o = p;
      if (a.materials[mat].type == NWPRT_MATERIAL_MIRROR) {
        d = reflect(rd, ns);
        if (!first_hit) mirror_first = true;
      } else {
        float2 tc = make_float2(0.0f, 0.0f);
        if (a.mat_tex) {
          tc = hit_texcoord(a.texcoords, a.indices, h.prim, h.u, h.v);
          w.tc[pair * slots + nrec] = tc;
        }
        const Surface sf = tm_surface(a, mat, ns, ng, tc);
        const float3 v = scale(rd, -1.0f);
        TmRecord &r = rec[nrec++];
        r.p[0] = p.x, r.p[1] = p.y, r.p[2] = p.z;
        r.v[0] = v.x, r.v[1] = v.y, r.v[2] = v.z;
        r.ns[0] = ns.x, r.ns[1] = ns.y, r.ns[2] = ns.z;
        r.ng[0] = ng.x, r.ng[1] = ng.y, r.ng[2] = ng.z;
        r.lc[0] = lc.x, r.lc[1] = lc.y, r.lc[2] = lc.z;
        r.tag = mat << 4 | depth;
        if (a.max_bounce == 1) {  // direct mode: BSDF light sample (devicePrograms.cu:202-247)
          TmBsdfLight &bl = *reinterpret_cast<TmBsdfLight *>(rec + 1);
          const float3 l = tm_bsdf_sample<DS>(sf, v, a.seed, vtx, c + 3);
          bl.visible = 0;
          if (l.x != -2.0f) {
            const float3 f = bsdf_eval<DS>(sf, l, v);
            bl.l[0] = l.x, bl.l[1] = l.y, bl.l[2] = l.z;
            bl.pb = bsdf_pdf<DS>(sf, l, v);
            bl.f[0] = f.x, bl.f[1] = f.y, bl.f[2] = f.z;
            bl.visible = !occluded(a.nodes, a.tris, p, l, stack);
          }
        }
        const float3 nd = tm_bsdf_sample<DS>(sf, v, a.seed, vtx, c + 6);
        if (nd.x == -2.0f) {
          done = true;
        } else {
          pdf = bsdf_pdf<DS>(sf, nd, v);
          if (pdf <= 1e-4f) {
            done = true;
          } else {
            const float3 f = bsdf_eval<DS>(sf, nd, v);
            d = nd;
            lc = v3(lc.x * f.x / pdf, lc.y * f.y / pdf, lc.z * f.z / pdf);
          }
        }
        first_hit = true;
      }
    }
    ++depth;
    if (!done && depth > 3) {  // Russian roulette (cameraShTmatrix.cu:139-151)
      const float z = rng_uniform(a.seed, vtx, c + 9);
      const float rr = fminf(fmaxf(lc.x, fmaxf(lc.y, lc.z)), 0.95f);
      if (z < rr) {
        const float q = fmaxf(rr, 1e-10f);
        lc = v3(lc.x / q, lc.y / q, lc.z / q);
      } else {
        done = true;
      }
    }
    if (done || depth >= a.max_bounce) break;
  }
  out.info = nrec | depth << 4 | (left ? 1 << 9 : 0) | (mirror_first ? 1 << 10 : 0);
  w.path[pair] = out;
  // Filtered basis 1 / (1 + 0.005 l^2 (l+1)^2) Y_k of the start direction
  // (cameraShTmatrix.cu:159-166).
  float y[kShCoeffs];
  sh_eval(tm_sphere(rng_uniform(a.seed, vtx, base), rng_uniform(a.seed, vtx, base + 1)), y);
  float4 *yo = reinterpret_cast<float4 *>(w.y + pair * kTmRow);
#pragma unroll
  for (int k4 = 0; k4 < kTmRow / 4; ++k4) {
    float q[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 4 * k4 + e;
      if (k < kShCoeffs) {
        int l = 0;
        while ((l + 1) * (l + 1) <= k) ++l;
        const float fl = (float)l;
        q[e] = 1.0f / (1.0f + 0.005f * fl * fl * (fl + 1.0f) * (fl + 1.0f)) * y[k];
      } else {
        q[e] = 0.0f;
      }
    }
    yo[k4] = make_float4(q[0], q[1], q[2], q[3]);
  }
}

__device__ inline bool tm_nan3(float3 c) { return isnan(c.x) || isnan(c.y) || isnan(c.z); }

template <bool DS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void tm_light_kernel(TmArgs a) {
  __shared__ int stack_lds[kBvhMaxDepth * 64];
  const int lane = threadIdx.x;
  const int nc = (a.degree + 1) * (a.degree + 1);
  const long long q = (long long)blockIdx.x * 64 + lane;
  const long long pairs = a.n_verts * a.samples;
  if (q >= pairs * nc) return;  // no barriers below
  int *stack = stack_lds + lane;
  const int slots = tm_slots(a.max_bounce);
  const TmWork w = tm_work(a.work, pairs, slots, nc);
  const long long pair = q / nc;
  const int i = (int)(q - pair * nc);
  const long long vl = pair / a.samples;
  const int s = (int)(pair - vl * a.samples);
  const uint32_t vtx = (uint32_t)(a.v_begin + vl);
  const uint32_t base = (uint32_t)s * kTmStride;
  const float *faces = a.basis + (size_t)i * 6 * a.n * a.n * 3;
  const double *tab = a.tabs + (size_t)i * a.tab_doubles;
  const TmPath path = w.path[pair];
  const int nrec = path.info & 15, traces = (path.info >> 4) & 31;
  const bool left = (path.info >> 9) & 1, mirror_first = (path.info >> 10) & 1;
  const TmRecord *rec = w.rec + pair * slots;
  float3 color = v3(0.0f, 0.0f, 0.0f);
  for (int r = 0; r < nrec; ++r) {
    const TmRecord h = rec[r];
    const uint32_t c = base + 16u * (uint32_t)((h.tag & 15) + 1);
    const float3 p = v3(h.p), v = v3(h.v), lc = v3(h.lc);
    const float3 l = env_sample(tab, a.n, rng_uniform(a.seed, vtx, c),
                                rng_uniform(a.seed, vtx, c + 1), rng_uniform(a.seed, vtx, c + 2));
    const float pe = env_pdf(tab, a.n, l);
    if (pe > 1e-4f) {
      // The term is formed before its shadow ray is cast: a zero term (a
      // direction below the surface, where the BSDF is zero) or a NaN one
      // adds nothing whether or not the ray is blocked, so only a non-zero
      // finite term pays for the ray.
      const Surface sf = tm_surface(a, h.tag >> 4, v3(h.ns), v3(h.ng),
                                    a.mat_tex ? w.tc[pair * slots + r] : make_float2(0.0f, 0.0f));
      const float pb = bsdf_pdf<DS>(sf, l, v);
      const float3 f = bsdf_eval<DS>(sf, l, v);
      const float3 le = env_radiance(faces, a.n, l);
      const float wt = pe / fmaxf(pe * pe + pb * pb, 1e-8f);
      const float3 inc = v3(f.x * le.x * lc.x * wt, f.y * le.y * lc.y * wt, f.z * le.z * lc.z * wt);
      if (!tm_nan3(inc) && (inc.x != 0.0f || inc.y != 0.0f || inc.z != 0.0f) &&
          !occluded(a.nodes, a.tris, p, l, stack))
        color = add(color, inc);
    }
    if (a.max_bounce == 1) {
      const TmBsdfLight bl = *reinterpret_cast<const TmBsdfLight *>(rec + 1);
      if (bl.visible) {
        const float3 l2 = v3(bl.l);
        const float pe2 = env_pdf(tab, a.n, l2);
        if (pe2 > 1e-4f) {
          const float3 le = env_radiance(faces, a.n, l2);
          const float wt = bl.pb / fmaxf(pe2 * pe2 + bl.pb * bl.pb, 1e-8f);
          const float3 inc = v3(bl.f[0] * le.x * lc.x * wt, bl.f[1] * le.y * lc.y * wt,
                                bl.f[2] * le.z * lc.z * wt);
          if (!tm_nan3(inc)) color = add(color, inc);
        }
      }
    }
  }
  if (left) {  // __miss__radiance (devicePrograms.cu:517-541)
    const float3 d = v3(path.dir), lc = v3(path.lc);
    const float3 le = env_radiance(faces, a.n, d);
    if (traces == 1) {
      color = le;
    } else if (path.pdf < 0.0f) {
      color = add(v3(le.x * lc.x, le.y * lc.y, le.z * lc.z), color);
    } else {
      const float pe = env_pdf(tab, a.n, d);
      const float wt = path.pdf * path.pdf / fmaxf(path.pdf * path.pdf + pe * pe, 1e-14f);
      color = add(v3(le.x * lc.x * wt, le.y * lc.y * wt, le.z * lc.z * wt), color);
    }
  }
  // The direct part: the first ray, or the ray after a first mirror, that
  // left the scene (cameraShTmatrix.cu:126-136).
  float3 tent = v3(0.0f, 0.0f, 0.0f);
  if (left && !tm_nan3(color) && (traces == 1 || (traces == 2 && mirror_first))) tent = color;
  float *o = w.ind + q * 3;
  o[0] = color.x - tent.x;
  o[1] = color.y - tent.y;
  o[2] = color.z - tent.z;
}

// The projection stages the filtered basis of 64 samples at a time in LDS;
// lane (map, channel) then walks them in sample order.
constexpr int kTmProjectSamples = 64;

__global__ __launch_bounds__(256) void tm_project_kernel(TmArgs a) {
  __shared__ float ylds[kTmProjectSamples][kTmRow];
  const int t = threadIdx.x;
  const int nc = (a.degree + 1) * (a.degree + 1), cols = 3 * nc;
  const long long vl = blockIdx.x;
  const long long pairs = a.n_verts * a.samples;
  const TmWork w = tm_work(a.work, pairs, tm_slots(a.max_bounce), nc);
  const float *ind = w.ind + (size_t)vl * a.samples * cols;
  const float fn = (float)a.samples, scale_n = 4.0f * 3.14159265358979f / fn;
  float acc[kShCoeffs];
#pragma unroll
  for (int k = 0; k < kShCoeffs; ++k) acc[k] = 0.0f;
  const int i = t / 3;
  for (int s0 = 0; s0 < a.samples; s0 += kTmProjectSamples) {
    const int m = min(kTmProjectSamples, a.samples - s0);
    __syncthreads();
    const float *ysrc = w.y + ((size_t)vl * a.samples + s0) * kTmRow;
    for (int e = t; e < m * kTmRow; e += 256) (&ylds[0][0])[e] = ysrc[e];
    __syncthreads();
    if (t < cols) {
      for (int j = 0; j < m; ++j) {
        const float *row = ind + (size_t)(s0 + j) * cols;
        const float x = row[t] * scale_n;
        if (isnan(row[3 * i]) || isnan(row[3 * i + 1]) || isnan(row[3 * i + 2])) {
#pragma unroll
          for (int k = 0; k < kShCoeffs; ++k)
            if (k < nc) acc[k] += acc[k] / fn;
        } else {
#pragma unroll
          for (int k = 0; k < kShCoeffs; ++k)
            if (k < nc) acc[k] = fmaf(ylds[j][k], x, acc[k]);
        }
      }
    }
  }
  if (t < cols) {
    float *o = a.transfer + (size_t)(a.v_begin + vl) * nc * cols;
#pragma unroll
    for (int k = 0; k < kShCoeffs; ++k)
      if (k < nc) o[(size_t)k * cols + t] = acc[k];
  }
}

}  // namespace

size_t env_table_doubles(int n) { return kTabHead + 6 * env_face_doubles(n); }

void launch_env_tables(const float *faces, int n, double *tab, hipStream_t s) {
  require(faces && tab && n >= 1 && n <= 128, NWPRT_ERR_INVALID_ARGUMENT,
          "env tables: face size must be in [1, 128]");
  hipLaunchKernelGGL(env_rows_kernel, dim3(6 * n), dim3(128), 0, s, faces, n, tab);
  hip_check(hipGetLastError(), "env_rows_kernel launch");
  hipLaunchKernelGGL(env_marginals_kernel, dim3(6), dim3(128), 0, s, n, tab);
  hip_check(hipGetLastError(), "env_marginals_kernel launch");
  hipLaunchKernelGGL(env_face_table_kernel, dim3(1), dim3(64), 0, s, tab);
  hip_check(hipGetLastError(), "env_face_table_kernel launch");
}

void launch_gbuffer_cast(const CastArgs &a, hipStream_t s) {
  require(a.nodes && a.tris && a.indices && a.materials, NWPRT_ERR_INVALID_ARGUMENT,
          "gbuffer_cast: mesh buffers missing");
  require(a.width > 0 && a.height > 0 && a.max_bounce >= 1, NWPRT_ERR_INVALID_ARGUMENT,
          "gbuffer_cast: bad size or bounce count");
  require(!a.env_faces || a.env_n > 0, NWPRT_ERR_INVALID_ARGUMENT,
          "gbuffer_cast: environment face size must be positive");
  require(!(a.gb.direct && a.light_faces) ||
              (a.env_tab && a.light_n >= 1 && a.light_n <= 128 && a.direct_samples >= 1),
          NWPRT_ERR_INVALID_ARGUMENT, "gbuffer_cast: bad direct-light arguments");
  const long long tiles = (long long)((a.width + 7) / 8) * ((a.height + 7) / 8);
  if (a.gb.direct && a.light_faces) {
    if (a.disney)
      hipLaunchKernelGGL((gbuffer_cast_kernel<true, true>), dim3((unsigned)tiles), dim3(64), 0, s, a);
    else
      hipLaunchKernelGGL((gbuffer_cast_kernel<true, false>), dim3((unsigned)tiles), dim3(64), 0, s, a);
  } else {
    hipLaunchKernelGGL((gbuffer_cast_kernel<false, false>), dim3((unsigned)tiles), dim3(64), 0, s, a);
  }
  hip_check(hipGetLastError(), "gbuffer_cast_kernel launch");
}

}  // namespace nwprt

namespace nwprt {

void launch_sh_render(const ShArgs &a, hipStream_t s) {
  require(a.nodes && a.tris && a.indices && a.vertices && a.normals && a.materials &&
              a.transfer && a.light_sh && a.light_response && a.corner_key && a.key_vertex &&
              a.key_material && a.n_keys > 0 && a.key_color && a.color,
          NWPRT_ERR_INVALID_ARGUMENT, "sh_render: null buffer");
  require(a.degree >= 0 && a.degree <= kShMaxDegree && a.samples >= 1 && a.width > 0 &&
              a.height > 0 && a.n_tris > 0 && a.n_vertices > 0,
          NWPRT_ERR_INVALID_ARGUMENT, "sh_render: bad degree, sample count or size");
  hipLaunchKernelGGL(sh_light_response_kernel, dim3((unsigned)((a.n_vertices + 3) / 4)),
                     dim3(256), 0, s, a);
  hip_check(hipGetLastError(), "sh_light_response_kernel launch");
  constexpr int pairs_per_block = 64 / kShPairLanes;
  const dim3 corner_grid((unsigned)((a.n_keys + pairs_per_block - 1) / pairs_per_block));
  if (a.disney)
    hipLaunchKernelGGL(sh_corner_kernel<true>, corner_grid, dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL(sh_corner_kernel<false>, corner_grid, dim3(64), 0, s, a);
  hip_check(hipGetLastError(), "sh_corner_kernel launch");
  const long long tiles = (long long)((a.width + 7) / 8) * ((a.height + 7) / 8);
  hipLaunchKernelGGL(sh_pixel_kernel, dim3((unsigned)tiles), dim3(64), 0, s, a);
  hip_check(hipGetLastError(), "sh_pixel_kernel launch");
}

size_t tm_work_bytes(long long n_verts, int samples, int degree, int max_bounce, bool textured) {
  const long long pairs = n_verts * samples;
  const int nc = (degree + 1) * (degree + 1);
  return tm_align((size_t)pairs * tm_slots(max_bounce) * sizeof(TmRecord)) +
         tm_align((size_t)pairs * sizeof(TmPath)) + tm_align((size_t)pairs * kTmRow * sizeof(float)) +
         tm_align((size_t)pairs * nc * 3 * sizeof(float)) +
         (textured ? tm_align((size_t)pairs * tm_slots(max_bounce) * sizeof(float2)) : 0);
}

void launch_sh_transfer(const TmArgs &a, hipStream_t s) {
  require(a.nodes && a.tris && a.indices && a.vertices && a.materials && a.basis && a.tabs &&
              a.work && a.transfer,
          NWPRT_ERR_INVALID_ARGUMENT, "sh_transfer: null buffer");
  require(a.degree >= 0 && a.degree <= kShMaxDegree && a.samples >= 1 && a.n >= 1 &&
              a.max_bounce >= 1 && a.max_bounce <= kTmMaxBounce && a.n_verts >= 1,
          NWPRT_ERR_INVALID_ARGUMENT, "sh_transfer: bad degree, samples, face size or bounces");
  const long long pairs = a.n_verts * a.samples;
  const int nc = (a.degree + 1) * (a.degree + 1);
  require(pairs * nc / 64 < (1ll << 31), NWPRT_ERR_INVALID_ARGUMENT,
          "sh_transfer: too many (vertex, sample, map) triples for one launch");
  const dim3 path_grid((unsigned)((pairs + 63) / 64)), light_grid((unsigned)((pairs * nc + 63) / 64));
  if (a.disney)
    hipLaunchKernelGGL(tm_paths_kernel<true>, path_grid, dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL(tm_paths_kernel<false>, path_grid, dim3(64), 0, s, a);
  hip_check(hipGetLastError(), "tm_paths_kernel launch");
  if (a.disney)
    hipLaunchKernelGGL(tm_light_kernel<true>, light_grid, dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL(tm_light_kernel<false>, light_grid, dim3(64), 0, s, a);
  hip_check(hipGetLastError(), "tm_light_kernel launch");
  hipLaunchKernelGGL(tm_project_kernel, dim3((unsigned)a.n_verts), dim3(256), 0, s, a);
  hip_check(hipGetLastError(), "tm_project_kernel launch");
}

}  // namespace nwprt
