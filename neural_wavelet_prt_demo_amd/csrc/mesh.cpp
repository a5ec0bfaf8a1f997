// Triangle meshes and the host side of the G-buffer producer and the SH-PRT
// render: BVH construction, upload, camera set-up and the nwprt_mesh_* /
// nwprt_gbuffer_cast / nwprt_sh_render entry points.  The reference builds an OptiX acceleration structure and launches
// its raygen program (OptixBackend::render, Renderer/OptixBackend.cpp:821-839);
// here the BVH is built on the host and traversed by gbuffer_kernels.hip.
//
// BVH: binned surface-area heuristic (16 bins per axis over the triangle
// centroids, traversal cost 1 against 1 per triangle test), leaves of at most
// 8 triangles, split at the centroid median wherever the heuristic's split
// would need more than kBvhMaxDepth internal levels below the root, so the
// device traversal stack never exceeds kBvhMaxDepth entries.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "engine_state.hpp"

using namespace nwprt;

namespace {

struct Box {
  float lo[3], hi[3];
  Box() { reset(); }
  void reset() {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::numeric_limits<float>::infinity();
      hi[a] = -std::numeric_limits<float>::infinity();
    }
  }
  void grow(const Box &b) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], b.lo[a]);
      hi[a] = std::max(hi[a], b.hi[a]);
    }
  }
  void grow(const float *p) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], p[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
  }
  double area() const {
    const double x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
    if (!(x >= 0 && y >= 0 && z >= 0)) return 0.0;
    return 2.0 * (x * y + y * z + z * x);
  }
};

// Internal levels a median-split subtree of n triangles needs.
int levels_needed(int64_t n) {
  int l = 0;
  while (n > kBvhLeafMax) {
    n = (n + 1) / 2;
    ++l;
  }
  return l;
}

constexpr int kBins = 16;
constexpr int kLeafTarget = 4;

struct Builder {
  std::vector<Box> tri_box;
  std::vector<float> cen;  // [n][3]
  std::vector<int32_t> order;
  std::vector<BvhNode> nodes;
  int max_depth = 0;

  static int32_t leaf_ref(int64_t first, int64_t count) {
    return ~(int32_t)(first * 8 + (count - 1));
  }

  Box range_box(int64_t b, int64_t e) const {
    Box bx;
    for (int64_t i = b; i < e; ++i) bx.grow(tri_box[order[i]]);
    return bx;
  }

  // Split position of [b, e) chosen by the binned SAH, or -1 when a leaf is
  // cheaper (and allowed) or no bin boundary separates the centroids.
  int64_t sah_split(int64_t b, int64_t e, const Box &bx) {
    const int64_t n = e - b;
    Box cb;
    for (int64_t i = b; i < e; ++i) cb.grow(&cen[3 * (size_t)order[i]]);
    double best = std::numeric_limits<double>::infinity();
    int best_axis = -1, best_bin = -1;
    for (int a = 0; a < 3; ++a) {
      const float ext = cb.hi[a] - cb.lo[a];
      if (!(ext > 0.0f)) continue;
      const float k = (float)kBins / ext;
      Box bins[kBins];
      int64_t cnt[kBins] = {};
      for (int64_t i = b; i < e; ++i) {
        const int32_t t = order[i];
        const int bi = std::min(kBins - 1, (int)((cen[3 * (size_t)t + a] - cb.lo[a]) * k));
        ++cnt[bi];
        bins[bi].grow(tri_box[t]);
      }
      double left_area[kBins];
      int64_t left_n[kBins];
      Box acc;
      int64_t an = 0;
      for (int j = 0; j < kBins; ++j) {
        acc.grow(bins[j]);
        an += cnt[j];
        left_area[j] = acc.area();
        left_n[j] = an;
      }
      acc.reset();
      an = 0;
      for (int j = kBins - 1; j > 0; --j) {
        acc.grow(bins[j]);
        an += cnt[j];
        const int64_t ln = left_n[j - 1];
        if (ln == 0 || an == 0) continue;
        const double c = left_area[j - 1] * (double)ln + acc.area() * (double)an;
        if (c < best) {
          best = c;
          best_axis = a;
          best_bin = j - 1;
        }
      }
    }
    if (best_axis < 0) return -1;
    const double pa = bx.area();
    const double split_cost = 1.0 + (pa > 0 ? best / pa : (double)n);
    if (n <= kBvhLeafMax && (double)n <= split_cost) return -1;
    const int a = best_axis;
    const float lo = cb.lo[a], k = (float)kBins / (cb.hi[a] - cb.lo[a]);
    auto it = std::partition(order.begin() + b, order.begin() + e, [&](int32_t t) {
      return std::min(kBins - 1, (int)((cen[3 * (size_t)t + a] - lo) * k)) <= best_bin;
    });
    const int64_t mid = it - order.begin();
    return (mid > b && mid < e) ? mid : -1;
  }

  // Median split of [b, e) along the widest centroid axis.
  int64_t median_split(int64_t b, int64_t e) {
    Box cb;
    for (int64_t i = b; i < e; ++i) cb.grow(&cen[3 * (size_t)order[i]]);
    int a = 0;
    for (int k = 1; k < 3; ++k)
      if (cb.hi[k] - cb.lo[k] > cb.hi[a] - cb.lo[a]) a = k;
    const int64_t mid = b + (e - b) / 2;
    std::nth_element(order.begin() + b, order.begin() + mid, order.begin() + e,
                     [&](int32_t x, int32_t y) {
                       const float cx = cen[3 * (size_t)x + a], cy = cen[3 * (size_t)y + a];
                       return cx < cy || (cx == cy && x < y);
                     });
    return mid;
  }

  // Reference to a child covering [b, e) whose node (if internal) sits at
  // internal level `level`; its box is returned in bx.
  int32_t child(int64_t b, int64_t e, int level, Box &bx) {
    bx = range_box(b, e);
    const int64_t n = e - b;
    if (n <= kLeafTarget || (n <= kBvhLeafMax && level >= kBvhMaxDepth)) return leaf_ref(b, n);
    int64_t mid = sah_split(b, e, bx);
    if (mid < 0 && n <= kBvhLeafMax) return leaf_ref(b, n);
    const int room = kBvhMaxDepth - (level + 1);
    if (mid < 0 || levels_needed(mid - b) > room || levels_needed(e - mid) > room)
      mid = median_split(b, e);
    return node(b, mid, e, level);
  }

  int32_t node(int64_t b, int64_t mid, int64_t e, int level) {
    require(level < kBvhMaxDepth, NWPRT_ERR_UNSUPPORTED, "mesh: BVH depth bound exceeded");
    max_depth = std::max(max_depth, level + 1);
    const size_t id = nodes.size();
    nodes.emplace_back();
    Box l, r;
    const int32_t cl = child(b, mid, level + 1, l);
    const int32_t cr = child(mid, e, level + 1, r);
    BvhNode &nd = nodes[id];
    std::memset(&nd, 0, sizeof nd);
    std::memcpy(nd.lo0, l.lo, sizeof l.lo);
    std::memcpy(nd.hi0, l.hi, sizeof l.hi);
    std::memcpy(nd.lo1, r.lo, sizeof r.lo);
    std::memcpy(nd.hi1, r.hi, sizeof r.hi);
    nd.c0 = cl;
    nd.c1 = cr;
    return (int32_t)id;
  }

  void build(int64_t n) {
    order.resize(n);
    std::iota(order.begin(), order.end(), 0);
    if (n <= kBvhLeafMax) {
      // Root whose two children are the same leaf (a second test of its few
      // triangles changes no hit: equal t and index never replace the hit).
      max_depth = 1;
      nodes.emplace_back();
      BvhNode &nd = nodes[0];
      std::memset(&nd, 0, sizeof nd);
      const Box bx = range_box(0, n);
      std::memcpy(nd.lo0, bx.lo, sizeof bx.lo);
      std::memcpy(nd.hi0, bx.hi, sizeof bx.hi);
      std::memcpy(nd.lo1, bx.lo, sizeof bx.lo);
      std::memcpy(nd.hi1, bx.hi, sizeof bx.hi);
      nd.c0 = nd.c1 = leaf_ref(0, n);
      return;
    }
    require(levels_needed(n) <= kBvhMaxDepth, NWPRT_ERR_UNSUPPORTED, "mesh: too many triangles");
    const int64_t mid0 = sah_split(0, n, range_box(0, n));
    const int64_t mid = (mid0 < 0 || levels_needed(mid0) > kBvhMaxDepth - 1 ||
                         levels_needed(n - mid0) > kBvhMaxDepth - 1)
                            ? median_split(0, n)
                            : mid0;
    node(0, mid, n, 0);
  }
};

Mesh &get_mesh(nwprt_ctx ctx, int id) {
  auto it = ctx->meshes.find(id);
  require(it != ctx->meshes.end(), NWPRT_ERR_NOT_FOUND, "unknown mesh id " + std::to_string(id));
  return *it->second;
}

// Drops the mesh's hold on its textures and its texture buffers.
void release_textures(nwprt_ctx ctx, Mesh &m) {
  for (int32_t id : m.texture_ids) {
    auto it = ctx->textures.find(id);
    if (it != ctx->textures.end()) --it->second->uses;
  }
  m.texture_ids.clear();
  m.texcoords.release();
  m.mat_tex.release();
  m.textured = false;
}

bool finite3(const float *p) {
  return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]);
}

// RenderEngine::setCamera (Renderer/RenderEngine.cpp:246-267), in double:
// direction, horizontal = aspect x unit(direction x up), vertical =
// unit(horizontal x direction), principal point at the image centre and a
// 0.8576 rad field of view.
void camera_basis(const nwprt_camera &c, int width, int height, float *pos, float *dir_out,
                  float *h_out, float *v_out, float &cx, float &cy, float &fx, float &fy) {
  double dir[3], up[3], h[3], v[3];
  for (int a = 0; a < 3; ++a) {
    dir[a] = (double)c.at[a] - c.from[a];
    up[a] = c.up[a];
  }
  auto norm = [](double *x) {
    const double l = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    require(l > 0 && std::isfinite(l), NWPRT_ERR_INVALID_ARGUMENT,
            "degenerate camera (from == at, or up parallel to the view)");
    for (int a = 0; a < 3; ++a) x[a] /= l;
  };
  auto cross = [](const double *x, const double *y, double *z) {
    z[0] = x[1] * y[2] - x[2] * y[1];
    z[1] = x[2] * y[0] - x[0] * y[2];
    z[2] = x[0] * y[1] - x[1] * y[0];
  };
  norm(dir);
  cross(dir, up, h);
  norm(h);
  const double aspect = (double)width / (double)height;
  for (int a = 0; a < 3; ++a) h[a] *= aspect;
  cross(h, dir, v);
  norm(v);
  for (int k = 0; k < 3; ++k) {
    pos[k] = c.from[k];
    dir_out[k] = (float)dir[k];
    h_out[k] = (float)h[k];
    v_out[k] = (float)v[k];
  }
  const double tan_half = std::tan(0.5 * 0.8575560548920328);
  cx = (float)(width / 2.0);
  cy = (float)(height / 2.0);
  fx = (float)(0.5 * width / tan_half);
  fy = (float)(0.5 * height / tan_half);
}

}  // namespace

extern "C" {

nwprt_status nwprt_mesh_create(nwprt_ctx ctx, const float *vertices, int64_t n_vertices,
                               const float *normals, const int32_t *indices,
                               int64_t n_triangles, const int32_t *material_ids,
                               const nwprt_material *materials, int32_t n_materials,
                               int32_t *mesh_id) {
  return guarded([&] {
    bind(ctx);
    require(vertices && indices && materials && mesh_id, NWPRT_ERR_INVALID_ARGUMENT,
            "mesh_create: null argument");
    require(n_vertices >= 3 && n_vertices <= INT32_MAX, NWPRT_ERR_INVALID_ARGUMENT,
            "mesh_create: vertex count must be in [3, 2^31)");
    require(n_triangles >= 1 && n_triangles < (1ll << 28), NWPRT_ERR_INVALID_ARGUMENT,
            "mesh_create: triangle count must be in [1, 2^28)");
    require(n_materials >= 1, NWPRT_ERR_INVALID_ARGUMENT, "mesh_create: no materials");
    for (int64_t i = 0; i < n_vertices; ++i)
      require(finite3(vertices + 3 * i) && (!normals || finite3(normals + 3 * i)),
              NWPRT_ERR_INVALID_ARGUMENT, "mesh_create: non-finite vertex or normal");
    for (int64_t i = 0; i < 3 * n_triangles; ++i)
      require(indices[i] >= 0 && indices[i] < n_vertices, NWPRT_ERR_INVALID_ARGUMENT,
              "mesh_create: vertex index out of range");
    if (material_ids)
      for (int64_t i = 0; i < n_triangles; ++i)
        require(material_ids[i] >= 0 && material_ids[i] < n_materials,
                NWPRT_ERR_INVALID_ARGUMENT, "mesh_create: material id out of range");
    std::vector<MaterialRec> mats(n_materials);
    for (int32_t i = 0; i < n_materials; ++i) {
      const nwprt_material &m = materials[i];
      const float params[11] = {m.specular_transmission, m.metallic, m.subsurface, m.specular,
                                m.specular_tint, m.anisotropic, m.sheen, m.sheen_tint,
                                m.clearcoat, m.clearcoat_gloss, m.eta};
      bool finite = finite3(m.kd) && finite3(m.ks) && std::isfinite(m.roughness);
      for (float x : params) finite = finite && std::isfinite(x);
      require(finite && m.type >= 0 && m.type <= 4 && m.eta > 0.0f,
              NWPRT_ERR_INVALID_ARGUMENT,
              "mesh_create: bad material " + std::to_string(i) +
                  " (non-finite value, type outside 0..4 or eta <= 0)");
      std::memset(&mats[i], 0, sizeof(MaterialRec));
      std::memcpy(mats[i].kd, m.kd, sizeof m.kd);
      std::memcpy(mats[i].ks, m.ks, sizeof m.ks);
      mats[i].roughness = m.roughness;
      mats[i].type = m.type;
      mats[i].specular_transmission = m.specular_transmission;
      mats[i].metallic = m.metallic;
      mats[i].subsurface = m.subsurface;
      mats[i].specular = m.specular;
      mats[i].specular_tint = m.specular_tint;
      mats[i].anisotropic = m.anisotropic;
      mats[i].sheen = m.sheen;
      mats[i].sheen_tint = m.sheen_tint;
      mats[i].clearcoat = m.clearcoat;
      mats[i].clearcoat_gloss = m.clearcoat_gloss;
      mats[i].eta = m.eta;
    }

    auto mesh = std::make_unique<Mesh>();
    Builder bld;
    bld.tri_box.resize(n_triangles);
    bld.cen.resize(3 * (size_t)n_triangles);
    Box all;
    for (int64_t i = 0; i < n_vertices; ++i) all.grow(vertices + 3 * i);
    for (int64_t t = 0; t < n_triangles; ++t) {
      Box &bx = bld.tri_box[t];
      for (int c = 0; c < 3; ++c) bx.grow(vertices + 3 * (size_t)indices[3 * t + c]);
      for (int a = 0; a < 3; ++a) bld.cen[3 * t + a] = 0.5f * (bx.lo[a] + bx.hi[a]);
    }
    bld.build(n_triangles);

    std::vector<BvhTri> tris(n_triangles);
    for (int64_t i = 0; i < n_triangles; ++i) {
      const int32_t t = bld.order[i];
      const float *a = vertices + 3 * (size_t)indices[3 * t];
      const float *b = vertices + 3 * (size_t)indices[3 * t + 1];
      const float *c = vertices + 3 * (size_t)indices[3 * t + 2];
      BvhTri &r = tris[i];
      std::memset(&r, 0, sizeof r);
      for (int k = 0; k < 3; ++k) {
        r.v0[k] = a[k];
        r.e1[k] = b[k] - a[k];
        r.e2[k] = c[k] - a[k];
      }
      r.prim = t;
    }
    mesh->nodes.upload(bld.nodes.data(), bld.nodes.size() * sizeof(BvhNode));
    mesh->tris.upload(tris.data(), tris.size() * sizeof(BvhTri));
    mesh->indices.upload(indices, (size_t)n_triangles * 3 * sizeof(int32_t));
    mesh->vertices.upload(vertices, (size_t)n_vertices * 3 * sizeof(float));
    if (normals) mesh->normals.upload(normals, (size_t)n_vertices * 3 * sizeof(float));
    if (material_ids)
      mesh->tri_material.upload(material_ids, (size_t)n_triangles * sizeof(int32_t));
    mesh->materials.upload(mats.data(), mats.size() * sizeof(MaterialRec));
    {
      // Corners that share a vertex and a material get the same SH colour:
      // number the distinct (vertex, material) pairs once, here.
      const int64_t n_corners = n_triangles * 3;
      std::vector<uint64_t> pair(n_corners);
      for (int64_t c = 0; c < n_corners; ++c) {
        const uint32_t mid = material_ids ? (uint32_t)material_ids[c / 3] : 0u;
        pair[c] = ((uint64_t)(uint32_t)indices[c] << 32) | mid;
      }
      std::vector<uint64_t> keys(pair);
      std::sort(keys.begin(), keys.end());
      keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
      std::vector<int32_t> corner_key(n_corners), key_vertex(keys.size()),
          key_material(keys.size());
      for (int64_t c = 0; c < n_corners; ++c)
        corner_key[c] =
            (int32_t)(std::lower_bound(keys.begin(), keys.end(), pair[c]) - keys.begin());
      for (size_t u = 0; u < keys.size(); ++u) {
        key_vertex[u] = (int32_t)(keys[u] >> 32);
        key_material[u] = (int32_t)(keys[u] & 0xffffffffu);
      }
      mesh->corner_key.upload(corner_key.data(), corner_key.size() * sizeof(int32_t));
      mesh->key_vertex.upload(key_vertex.data(), key_vertex.size() * sizeof(int32_t));
      mesh->key_material.upload(key_material.data(), key_material.size() * sizeof(int32_t));
      mesh->n_keys = (int64_t)keys.size();
    }
    std::memcpy(mesh->lower, all.lo, sizeof all.lo);
    std::memcpy(mesh->upper, all.hi, sizeof all.hi);
    mesh->n_tris = n_triangles;
    mesh->n_vertices = n_vertices;
    mesh->n_nodes = (int64_t)bld.nodes.size();
    mesh->depth = bld.max_depth;
    mesh->has_normals = normals != nullptr;
    mesh->has_tri_material = material_ids != nullptr;
    mesh->n_materials = n_materials;
    // Capabilities of the materials the triangles use.
    std::vector<char> used(mats.size(), 0);
    if (material_ids)
      for (int64_t i = 0; i < n_triangles; ++i) used[(size_t)material_ids[i]] = 1;
    else
      used[0] = 1;
    for (size_t i = 0; i < mats.size(); ++i) {
      if (!used[i]) continue;
      if (mats[i].type == NWPRT_MATERIAL_MIRROR) mesh->sh_ok = false;
      if (mats[i].type == 1) mesh->disney = true;
    }
    const int id = ctx->next_id++;
    ctx->meshes[id] = std::move(mesh);
    *mesh_id = id;
  });
}

nwprt_status nwprt_mesh_get_info(nwprt_ctx ctx, int32_t mesh_id, nwprt_mesh_info *out) {
  return guarded([&] {
    require(ctx && out, NWPRT_ERR_INVALID_ARGUMENT, "mesh_get_info: null argument");
    const Mesh &m = get_mesh(ctx, mesh_id);
    std::memcpy(out->lower, m.lower, sizeof m.lower);
    std::memcpy(out->upper, m.upper, sizeof m.upper);
    out->triangles = m.n_tris;
    out->bvh_nodes = m.n_nodes;
    out->bvh_depth = m.depth;
    out->vertices = m.n_vertices;
  });
}

nwprt_status nwprt_mesh_destroy(nwprt_ctx ctx, int32_t mesh_id) {
  return guarded([&] {
    bind(ctx);
    Mesh &m = get_mesh(ctx, mesh_id);
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    release_textures(ctx, m);
    ctx->meshes.erase(mesh_id);
  });
}

nwprt_status nwprt_texture_create(nwprt_ctx ctx, const uint8_t *rgba, int32_t width,
                                  int32_t height, int32_t *texture_id) {
  return guarded([&] {
    bind(ctx);
    require(rgba && texture_id, NWPRT_ERR_INVALID_ARGUMENT, "texture_create: null argument");
    require(width >= 1 && height >= 1 && (int64_t)width * height <= (1ll << 28),
            NWPRT_ERR_INVALID_ARGUMENT, "texture_create: size must be 1..2^28 texels");
    auto t = std::make_unique<TextureObj>();
    t->rgba.upload(rgba, (size_t)width * height * 4);
    t->width = width;
    t->height = height;
    const int id = ctx->next_id++;
    ctx->textures[id] = std::move(t);
    *texture_id = id;
  });
}

nwprt_status nwprt_texture_destroy(nwprt_ctx ctx, int32_t texture_id) {
  return guarded([&] {
    bind(ctx);
    auto it = ctx->textures.find(texture_id);
    require(it != ctx->textures.end(), NWPRT_ERR_NOT_FOUND,
            "unknown texture id " + std::to_string(texture_id));
    require(it->second->uses == 0, NWPRT_ERR_STATE, "texture_destroy: a mesh uses the texture");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    ctx->textures.erase(it);
  });
}

nwprt_status nwprt_mesh_set_textures(nwprt_ctx ctx, int32_t mesh_id, const float *texcoords,
                                     const int32_t *material_textures) {
  return guarded([&] {
    bind(ctx);
    Mesh &m = get_mesh(ctx, mesh_id);
    std::vector<TexRef> refs;
    std::vector<int32_t> held;
    if (texcoords && material_textures) {
      refs.resize((size_t)m.n_materials * kTexFields, TexRef{nullptr, 0, 0});
      for (size_t i = 0; i < refs.size(); ++i) {
        const int32_t id = material_textures[i];
        if (id < 0) continue;
        auto it = ctx->textures.find(id);
        require(it != ctx->textures.end(), NWPRT_ERR_NOT_FOUND,
                "mesh_set_textures: unknown texture id " + std::to_string(id));
        refs[i] = TexRef{it->second->rgba.as<uint8_t>(), it->second->width, it->second->height};
        held.push_back(id);
      }
      for (int64_t v = 0; v < 2 * m.n_vertices; ++v)
        require(std::isfinite(texcoords[v]), NWPRT_ERR_INVALID_ARGUMENT,
                "mesh_set_textures: texture coordinates must be finite");
    }
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");  // kernels may read the old ones
    release_textures(ctx, m);
    if (held.empty()) return;  // no texture read: the mesh stays untextured
    m.texcoords.upload(texcoords, (size_t)m.n_vertices * 2 * sizeof(float));
    m.mat_tex.upload(refs.data(), refs.size() * sizeof(TexRef));
    for (int32_t id : held) ++ctx->textures[id]->uses;
    m.texture_ids = std::move(held);
    m.textured = true;
  });
}

nwprt_status nwprt_gbuffer_cast(nwprt_ctx ctx, int32_t mesh_id, const nwprt_camera *camera,
                                int32_t width, int32_t height,
                                const nwprt_cast_params *params,
                                const nwprt_gbuffer *gbuffer, void *stream) {
  return guarded([&] {
    bind(ctx);
    const Mesh &m = get_mesh(ctx, mesh_id);
    require(camera && params && gbuffer, NWPRT_ERR_INVALID_ARGUMENT,
            "gbuffer_cast: null argument");
    require(width > 0 && height > 0 && (int64_t)width * height <= (1ll << 31),
            NWPRT_ERR_INVALID_ARGUMENT, "gbuffer_cast: bad image size");
    require(params->max_bounce >= 1 && params->max_bounce <= 64, NWPRT_ERR_INVALID_ARGUMENT,
            "gbuffer_cast: max_bounce must be in [1, 64]");
    require(!params->view_env_faces || params->view_env_n > 0, NWPRT_ERR_INVALID_ARGUMENT,
            "gbuffer_cast: environment face size must be positive");
    require(finite3(camera->from) && finite3(camera->at) && finite3(camera->up),
            NWPRT_ERR_INVALID_ARGUMENT, "gbuffer_cast: non-finite camera");

    CastArgs a{};
    camera_basis(*camera, width, height, a.cam_pos, a.cam_dir, a.cam_h, a.cam_v, a.cx, a.cy, a.fx,
                 a.fy);
    a.nodes = m.nodes.as<BvhNode>();
    a.tris = m.tris.as<BvhTri>();
    a.indices = m.indices.as<int32_t>();
    a.normals = m.has_normals ? m.normals.as<float>() : nullptr;
    a.tri_material = m.has_tri_material ? m.tri_material.as<int32_t>() : nullptr;
    a.materials = m.materials.as<MaterialRec>();
    a.disney = m.disney;
    a.texcoords = m.textured ? m.texcoords.as<float>() : nullptr;
    a.mat_tex = m.textured ? m.mat_tex.as<TexRef>() : nullptr;
    const float *lo = params->bounds ? params->bounds : m.lower;
    const float *hi = params->bounds ? params->bounds + 3 : m.upper;
    for (int k = 0; k < 3; ++k) {
      require(std::isfinite(lo[k]) && std::isfinite(hi[k]) && hi[k] > lo[k],
              NWPRT_ERR_INVALID_ARGUMENT,
              "gbuffer_cast: bounds must have a positive extent on every axis "
              "(pass explicit bounds for a flat mesh)");
      a.lower[k] = lo[k];
      a.box[k] = hi[k] - lo[k];
    }
    a.width = width;
    a.height = height;
    a.max_bounce = params->max_bounce;
    a.env_faces = params->view_env_faces;
    a.env_n = params->view_env_n;
    a.gb = *gbuffer;
    if (params->light_env_faces && gbuffer->direct) {
      require(params->light_env_n >= 1 && params->light_env_n <= NWPRT_MAX_FACE_SIZE,
              NWPRT_ERR_INVALID_ARGUMENT, "gbuffer_cast: light face size must be in [1, 128]");
      require(params->direct_samples >= 1 && params->direct_samples <= (1 << 20),
              NWPRT_ERR_INVALID_ARGUMENT, "gbuffer_cast: direct_samples must be in [1, 2^20]");
      check_aligned(gbuffer->direct, 16, "G-buffer direct");
      ctx->env_tables.ensure(env_table_doubles(params->light_env_n) * sizeof(double));
      launch_env_tables(params->light_env_faces, params->light_env_n,
                        ctx->env_tables.as<double>(), pick(ctx, stream));
      a.light_faces = params->light_env_faces;
      a.env_tab = ctx->env_tables.as<double>();
      a.light_n = params->light_env_n;
      a.direct_samples = params->direct_samples;
      a.seed = params->seed;
    } else {
      a.gb.direct = nullptr;
    }
    for (const void *p : {(const void *)gbuffer->position, (const void *)gbuffer->normal,
                          (const void *)gbuffer->reflect, (const void *)gbuffer->kd,
                          (const void *)gbuffer->ks})
      check_aligned(p, 16, "G-buffer float4 field");
    check_aligned(gbuffer->roughness, 4, "G-buffer roughness");
    launch_gbuffer_cast(a, pick(ctx, stream));
  });
}

nwprt_status nwprt_sh_render(nwprt_ctx ctx, int32_t mesh_id, const nwprt_camera *camera,
                             int32_t width, int32_t height, const float *transfer,
                             const float *light_sh, int32_t degree, int32_t brdf_samples,
                             uint32_t seed, const float *view_env_faces, int32_t view_env_n,
                             float *color, void *stream) {
  return guarded([&] {
    bind(ctx);
    const Mesh &m = get_mesh(ctx, mesh_id);
    require(camera && transfer && light_sh && color, NWPRT_ERR_INVALID_ARGUMENT,
            "sh_render: null argument");
    require(width > 0 && height > 0 && (int64_t)width * height <= (1ll << 31),
            NWPRT_ERR_INVALID_ARGUMENT, "sh_render: bad image size");
    require(degree >= 0 && degree <= kShMaxDegree, NWPRT_ERR_INVALID_ARGUMENT,
            "sh_render: SH degree must be in [0, 8]");
    require(brdf_samples >= 1 && brdf_samples <= (1 << 20), NWPRT_ERR_INVALID_ARGUMENT,
            "sh_render: brdf_samples must be in [1, 2^20]");
    require(!view_env_faces || (view_env_n >= 1 && view_env_n <= NWPRT_MAX_FACE_SIZE),
            NWPRT_ERR_INVALID_ARGUMENT, "sh_render: bad view environment size");
    require(m.has_normals, NWPRT_ERR_INVALID_ARGUMENT,
            "sh_render: the mesh needs vertex normals");
    require(m.sh_ok, NWPRT_ERR_UNSUPPORTED,
            "sh_render: mirror materials have no SH-render BSDF");
    require(!m.textured, NWPRT_ERR_UNSUPPORTED,
            "sh_render: textured meshes are read per pixel by the reference; the per-vertex "
            "colours of this render cannot carry them");
    require(finite3(camera->from) && finite3(camera->at) && finite3(camera->up),
            NWPRT_ERR_INVALID_ARGUMENT, "sh_render: non-finite camera");
    check_aligned(color, 16, "sh_render colour");
    ShArgs a{};
    camera_basis(*camera, width, height, a.cam_pos, a.cam_dir, a.cam_h, a.cam_v, a.cx, a.cy, a.fx,
                 a.fy);
    const int nc = (degree + 1) * (degree + 1);
    ctx->sh_response.ensure((size_t)m.n_vertices * nc * 3 * sizeof(float));
    ctx->sh_colors.ensure((size_t)m.n_keys * 3 * sizeof(float));
    a.nodes = m.nodes.as<BvhNode>();
    a.tris = m.tris.as<BvhTri>();
    a.indices = m.indices.as<int32_t>();
    a.vertices = m.vertices.as<float>();
    a.normals = m.normals.as<float>();
    a.tri_material = m.has_tri_material ? m.tri_material.as<int32_t>() : nullptr;
    a.materials = m.materials.as<MaterialRec>();
    a.disney = m.disney;
    a.n_tris = m.n_tris;
    a.n_vertices = m.n_vertices;
    a.transfer = transfer;
    a.light_sh = light_sh;
    a.degree = degree;
    a.samples = brdf_samples;
    a.seed = seed;
    a.light_response = ctx->sh_response.as<float>();
    a.corner_key = m.corner_key.as<int32_t>();
    a.key_vertex = m.key_vertex.as<int32_t>();
    a.key_material = m.key_material.as<int32_t>();
    a.n_keys = m.n_keys;
    a.key_color = ctx->sh_colors.as<float>();
    a.width = width;
    a.height = height;
    a.env_faces = view_env_faces;
    a.env_n = view_env_n;
    a.color = color;
    launch_sh_render(a, pick(ctx, stream));
  });
}

nwprt_status nwprt_sh_compute_transfer(nwprt_ctx ctx, int32_t mesh_id, const float *basis_faces,
                                       int32_t face_n, int32_t degree, int32_t samples,
                                       int32_t max_bounce, uint32_t seed, int64_t vertex_begin,
                                       int64_t vertex_count, float *transfer, void *stream) {
  return guarded([&] {
    bind(ctx);
    const Mesh &m = get_mesh(ctx, mesh_id);
    require(basis_faces && transfer, NWPRT_ERR_INVALID_ARGUMENT, "sh_compute_transfer: null argument");
    require(degree >= 0 && degree <= kShMaxDegree, NWPRT_ERR_INVALID_ARGUMENT,
            "sh_compute_transfer: SH degree must be in [0, 8]");
    require(face_n >= 1 && face_n <= NWPRT_MAX_FACE_SIZE, NWPRT_ERR_INVALID_ARGUMENT,
            "sh_compute_transfer: face size must be in [1, 128]");
    require(samples >= 1 && samples <= (1 << 20), NWPRT_ERR_INVALID_ARGUMENT,
            "sh_compute_transfer: samples must be in [1, 2^20]");
    require(max_bounce >= 1 && max_bounce <= kTmMaxBounce, NWPRT_ERR_INVALID_ARGUMENT,
            "sh_compute_transfer: max_bounce must be in [1, 15]");
    if (vertex_count < 0) vertex_count = m.n_vertices - vertex_begin;
    require(vertex_begin >= 0 && vertex_count >= 0 && vertex_begin + vertex_count <= m.n_vertices,
            NWPRT_ERR_INVALID_ARGUMENT, "sh_compute_transfer: vertex range outside the mesh");
    check_aligned(basis_faces, 4, "basis faces");
    check_aligned(transfer, 4, "transfer");
    if (vertex_count == 0) return;
    const hipStream_t st = pick(ctx, stream);
    const int nc = (degree + 1) * (degree + 1);
    const size_t td = env_table_doubles(face_n), face_floats = (size_t)6 * face_n * face_n * 3;
    ctx->sh_tables.ensure(td * nc * sizeof(double));
    for (int i = 0; i < nc; ++i)
      launch_env_tables(basis_faces + face_floats * i, face_n, ctx->sh_tables.as<double>() + td * i,
                        st);
    // Vertices per launch: the path records, start-direction bases and
    // per-(sample, map) radiance of a launch stay within kTmWorkBudget, and
    // the (vertex, sample, map) triples within one grid.
    constexpr size_t kTmWorkBudget = (size_t)2 << 30;
    const size_t per_vertex = tm_work_bytes(1, samples, degree, max_bounce, m.textured);
    long long chunk = (long long)std::max<size_t>(1, kTmWorkBudget / per_vertex);
    chunk = std::min<long long>(chunk, ((1ll << 31) - 1) * 64 / ((long long)samples * nc));
    chunk = std::max<long long>(1, std::min<long long>(chunk, vertex_count));
    ctx->sh_work.ensure(tm_work_bytes(chunk, samples, degree, max_bounce, m.textured));
    TmArgs a{};
    a.nodes = m.nodes.as<BvhNode>();
    a.tris = m.tris.as<BvhTri>();
    a.indices = m.indices.as<int32_t>();
    a.vertices = m.vertices.as<float>();
    a.normals = m.has_normals ? m.normals.as<float>() : nullptr;
    a.tri_material = m.has_tri_material ? m.tri_material.as<int32_t>() : nullptr;
    a.materials = m.materials.as<MaterialRec>();
    a.disney = m.disney;
    a.texcoords = m.textured ? m.texcoords.as<float>() : nullptr;
    a.mat_tex = m.textured ? m.mat_tex.as<TexRef>() : nullptr;
    a.basis = basis_faces;
    a.tabs = ctx->sh_tables.as<double>();
    a.tab_doubles = td;
    a.n = face_n;
    a.degree = degree;
    a.samples = samples;
    a.max_bounce = max_bounce;
    a.seed = seed;
    a.work = ctx->sh_work.p;
    a.transfer = transfer;
    for (long long v0 = 0; v0 < vertex_count; v0 += chunk) {
      a.v_begin = vertex_begin + v0;
      a.n_verts = std::min<long long>(chunk, vertex_count - v0);
      launch_sh_transfer(a, st);
    }
  });
}

}  // extern "C"
