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

// One 8 x 8 pixel tile per 64-lane workgroup (coherent primary rays).
__global__ __launch_bounds__(64) void gbuffer_cast_kernel(CastArgs a) {
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
  float3 kd = nrm, ks = nrm;
  float rough = 0.0f;
  unsigned char flag = 0;
  for (int depth = 1;; ++depth) {
    const Hit h = trace(a.nodes, a.tris, o, d, stack);
    if (h.prim == INT_MAX) {
      nrm = scale(d, -1.0f);
      refl = nrm;
      pos = v3(-100000.0f, -100000.0f, -100000.0f);
      if (depth == 1 && a.env_faces) kd = env_radiance(a.env_faces, a.env_n, d);
      ks = v3(0.0f, 0.0f, 0.0f);
      rough = 1.0f;
      break;
    }
    const MaterialRec mat = a.materials[a.tri_material ? a.tri_material[h.prim] : 0];
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
    pos = p;
    nrm = ns;
    refl = normalize(reflect(d, ns));
    kd = v3(mat.kd);
    ks = v3(mat.ks);
    rough = mat.roughness;
    flag = 1;
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
}

}  // namespace

void launch_gbuffer_cast(const CastArgs &a, hipStream_t s) {
  require(a.nodes && a.tris && a.indices && a.materials, NWPRT_ERR_INVALID_ARGUMENT,
          "gbuffer_cast: mesh buffers missing");
  require(a.width > 0 && a.height > 0 && a.max_bounce >= 1, NWPRT_ERR_INVALID_ARGUMENT,
          "gbuffer_cast: bad size or bounce count");
  require(!a.env_faces || a.env_n > 0, NWPRT_ERR_INVALID_ARGUMENT,
          "gbuffer_cast: environment face size must be positive");
  const long long tiles = (long long)((a.width + 7) / 8) * ((a.height + 7) / 8);
  hipLaunchKernelGGL(gbuffer_cast_kernel, dim3((unsigned)tiles), dim3(64), 0, s, a);
  hip_check(hipGetLastError(), "gbuffer_cast_kernel launch");
}

}  // namespace nwprt
