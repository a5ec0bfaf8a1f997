// C ABI of the host-only utilities (no GPU use): level table, parameter
// counts, seeded inputs and file formats (npy, hdr, cube crosses, OBJ).
#include <array>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>

#include "nwprt_internal.hpp"

using namespace nwprt;

namespace {
template <class F>
nwprt_status host_guarded(F &&f) {
  try {
    f();
    set_last_error("");
    return NWPRT_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return NWPRT_ERR_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return NWPRT_ERR_INVALID_ARGUMENT;
  }
}

template <class T>
T *copy_out(const std::vector<T> &v) {
  T *p = (T *)std::malloc(v.size() * sizeof(T) + 1);
  if (!p) throw std::bad_alloc();
  if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

void shape_out(const NpyArray &a, int64_t *shape, int32_t *ndim, int32_t *fortran) {
  require(a.shape.size() <= 8, NWPRT_ERR_UNSUPPORTED, "npy: more than 8 dimensions");
  if (shape)
    for (size_t i = 0; i < a.shape.size(); ++i) shape[i] = a.shape[i];
  if (ndim) *ndim = (int32_t)a.shape.size();
  if (fortran) *fortran = a.fortran_order ? 1 : 0;
}
}  // namespace

extern "C" {

const char *nwprt_last_error(void) { return last_error_cstr(); }

void nwprt_free_host(void *p) { std::free(p); }

nwprt_status nwprt_hashgrid_level_table(const nwprt_hashgrid_desc *desc,
                                        nwprt_hashgrid_levels *out) {
  return host_guarded([&] {
    require(desc && out, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    hashgrid_levels(*desc, *out);
  });
}

nwprt_status nwprt_scene_param_counts(const nwprt_scene_desc *desc, size_t *hash_entries,
                                      size_t *wavelet_feature_count, size_t *projection_count,
                                      size_t *layer_weight_count) {
  return host_guarded([&] {
    require(desc != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    validate_scene_desc(*desc);
    if (hash_entries) *hash_entries = scene_hash_entries(*desc);
    if (wavelet_feature_count)
      *wavelet_feature_count = (size_t)desc->feature_dim * 6 * desc->face_size * desc->face_size;
    if (projection_count)
      *projection_count = desc->has_projection ? (size_t)desc->feature_dim * desc->feature_dim : 0;
    if (layer_weight_count) *layer_weight_count = scene_layer_weight_count(*desc);
  });
}

nwprt_status nwprt_synth_scene(const nwprt_scene_desc *desc, uint64_t seed, float hash_std,
                               float feature_std, void *hash_table, float *wavelet_features,
                               float *projection, float *layers) {
  return host_guarded([&] {
    require(desc != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    synth_scene(*desc, seed, hash_std, feature_std, hash_table, wavelet_features, projection,
                layers);
  });
}

nwprt_status nwprt_synth_gbuffer(uint64_t seed, int32_t width, int32_t height,
                                 float background_fraction, const nwprt_gbuffer *host_out) {
  return host_guarded([&] {
    require(host_out != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    synth_gbuffer(seed, width, height, background_fraction, *host_out);
  });
}

nwprt_status nwprt_synth_cubemap(uint64_t seed, int32_t face_size, float *faces) {
  return host_guarded([&] {
    require(faces != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    synth_cubemap(seed, face_size, faces);
  });
}

nwprt_status nwprt_npy_read_f32(const char *path, float **data, int64_t *shape, int32_t *ndim,
                                int32_t *fortran_order) {
  return host_guarded([&] {
    require(path && data, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    const NpyArray a = npy_read(path);
    shape_out(a, shape, ndim, fortran_order);
    *data = copy_out(a.as_f32());
  });
}

nwprt_status nwprt_npy_read_i32(const char *path, int32_t **data, int64_t *shape, int32_t *ndim,
                                int32_t *fortran_order) {
  return host_guarded([&] {
    require(path && data, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    const NpyArray a = npy_read(path);
    shape_out(a, shape, ndim, fortran_order);
    *data = copy_out(a.as_i32());
  });
}

nwprt_status nwprt_npy_write_f32(const char *path, const float *data, const int64_t *shape,
                                 int32_t ndim) {
  return host_guarded([&] {
    require(path && data && shape && ndim >= 0 && ndim <= 8, NWPRT_ERR_INVALID_ARGUMENT,
            "bad argument");
    npy_write(path, data, 'f', 4, std::vector<int64_t>(shape, shape + ndim));
  });
}

nwprt_status nwprt_npy_write_i32(const char *path, const int32_t *data, const int64_t *shape,
                                 int32_t ndim) {
  return host_guarded([&] {
    require(path && data && shape && ndim >= 0 && ndim <= 8, NWPRT_ERR_INVALID_ARGUMENT,
            "bad argument");
    npy_write(path, data, 'i', 4, std::vector<int64_t>(shape, shape + ndim));
  });
}

nwprt_status nwprt_hdr_read(const char *path, float **rgb, int32_t *width, int32_t *height) {
  return host_guarded([&] {
    require(path && rgb && width && height, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    const HdrImage img = hdr_read(path);
    *rgb = copy_out(img.rgb);
    *width = img.width;
    *height = img.height;
  });
}

nwprt_status nwprt_hdr_write(const char *path, const float *rgb, int32_t width, int32_t height,
                             int32_t channels) {
  return host_guarded([&] {
    require(path && rgb, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    hdr_write(path, rgb, width, height, channels);
  });
}

nwprt_status nwprt_cubemap_cross_to_faces(const float *cross, int32_t width, int32_t height,
                                          float *faces) {
  return host_guarded([&] {
    require(cross && faces, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    cross_to_faces(cross, width, height, faces);
  });
}

}  // extern "C"

// ------------------------------------------------------------------ OBJ --
// Wavefront OBJ + MTL reader for the G-buffer producer's meshes, following
// what loadOBJ (Renderer/Model.cpp:73-146, tinyobjloader with triangulation)
// keeps: positions, normals, texture coordinates, polygon faces fanned into
// triangles, and per-face materials made into cookTorrance materials (kd =
// Kd, ks = Ks, roughness = (1 / Ns)^2, 0 for Ns = 0; faces without a
// material get cookTorranceMaterial's defaults kd = ks = 0.5, roughness 1,
// Renderer/Material.h:136-143) with their map_Kd / map_Ks / map_Ns files
// (Model.cpp:133-135).  Vertices come in the reference's order: one mesh
// per (shape, material), shapes split at o / g lines, materials ascending,
// each mesh numbering its distinct (position, normal, texcoord) corners by
// first use (addVertex, Model.cpp:34-70, normals and texcoords back-filled
// the same way).  The reference keeps one corner map per shape across its
// materials' meshes, so a corner shared by two materials of one shape
// indexes the first mesh's vertices from the second; here each mesh keys
// its own corners.
namespace {

struct ObjMesh {
  std::vector<float> vertices, normals, texcoords;
  std::vector<int32_t> indices, material_ids;
  std::vector<nwprt_material> materials;
  // Per material: the kd, ks and roughness map files (map_Kd, map_Ks,
  // map_Ns), resolved against the OBJ's folder; empty for none.
  std::vector<std::string> maps;
  bool has_normals = false, has_texcoords = false;
};

}  // namespace

// Material::init (Renderer/Material.h:79-115) and the per-type constructors
// (:119-155).
nwprt_material material_defaults(int type) {
  nwprt_material m{};
  for (int c = 0; c < 3; ++c) m.kd[c] = m.ks[c] = 0.25f;
  m.roughness = 0.5f;
  m.type = type;
  m.specular_transmission = 0.f;
  m.metallic = 0.f;
  m.subsurface = 0.f;
  m.specular = 0.5f;
  m.specular_tint = 0.f;
  m.anisotropic = 0.f;
  m.sheen = 0.f;
  m.sheen_tint = 0.5f;
  m.clearcoat = 0.f;
  m.clearcoat_gloss = 1.f;
  m.eta = 1.49f / 1.000277f;
  switch (type) {
    case 0:  // roughPlasticMaterial
      for (int c = 0; c < 3; ++c) m.kd[c] = 0.5f, m.ks[c] = 1.0f;
      m.roughness = 0.1f;
      break;
    case 1:  // DisneyBSDFMaterial
      for (int c = 0; c < 3; ++c) m.kd[c] = 0.5f;
      m.roughness = 0.5f;
      break;
    case 2:  // cookTorranceMaterial
      for (int c = 0; c < 3; ++c) m.kd[c] = m.ks[c] = 0.5f;
      m.roughness = 1.0f;
      break;
    case 4:  // simple_ggxMaterial
      for (int c = 0; c < 3; ++c) m.kd[c] = 0.5f;
      m.roughness = 1.0f;
      break;
    default:
      break;
  }
  return m;
}

namespace {

nwprt_material cook_torrance(float kd0, float kd1, float kd2, float ks0, float ks1, float ks2,
                             float ns) {
  nwprt_material m = material_defaults(2);
  m.kd[0] = kd0, m.kd[1] = kd1, m.kd[2] = kd2;
  m.ks[0] = ks0, m.ks[1] = ks1, m.ks[2] = ks2;
  m.roughness = ns == 0.f ? 0.f : (float)std::pow(1.0 / ns, 2.0);
  m.type = 2;  // MatType::cookTorrance
  return m;
}

std::vector<std::string> split_ws(const std::string &line) {
  std::vector<std::string> out;
  std::istringstream ss(line);
  for (std::string w; ss >> w;) out.push_back(w);
  return out;
}

void read_mtl(const std::string &path, const std::string &dir, std::map<std::string, int> &names,
              std::vector<nwprt_material> &mats, std::vector<std::string> &maps) {
  std::ifstream in(path);
  require((bool)in, NWPRT_ERR_IO, "obj: cannot open material library " + path);
  int cur = -1;
  float ns = 0.f;
  auto close = [&] {
    if (cur >= 0) mats[(size_t)cur].roughness = ns == 0.f ? 0.f : (float)std::pow(1.0 / ns, 2.0);
  };
  for (std::string line; std::getline(in, line);) {
    const auto t = split_ws(line);
    if (t.empty() || t[0][0] == '#') continue;
    if (t[0] == "newmtl" && t.size() > 1) {
      close();
      names[t[1]] = cur = (int)mats.size();
      // tinyobjloader's material defaults: Kd = Ks = 0, Ns = 1.
      mats.push_back(cook_torrance(0, 0, 0, 0, 0, 0, 1.f));
      maps.resize(mats.size() * 3);
      ns = 1.f;
    } else if (cur >= 0 && (t[0] == "Kd" || t[0] == "Ks") && t.size() >= 4) {
      float *dst = t[0] == "Kd" ? mats[(size_t)cur].kd : mats[(size_t)cur].ks;
      for (int k = 0; k < 3; ++k) dst[k] = std::stof(t[(size_t)k + 1]);
    } else if (cur >= 0 && t[0] == "Ns" && t.size() >= 2) {
      ns = std::stof(t[1]);
    } else if (cur >= 0 && (t[0] == "map_Kd" || t[0] == "map_Ks" || t[0] == "map_Ns") &&
               t.size() >= 2) {
      // The file name is the last token (options such as -s come first).
      const int slot = t[0] == "map_Kd" ? 0 : t[0] == "map_Ks" ? 1 : 2;
      const std::string &f = t.back();
      maps[(size_t)cur * 3 + slot] = f.size() && f[0] == '/' ? f : dir + f;
    }
  }
  close();
}

struct Corner {
  int v, t, n;  // position, texcoord, normal (-1: none)
  bool operator<(const Corner &o) const {  // tinyobj::index_t order
    if (v != o.v) return v < o.v;
    if (n != o.n) return n < o.n;
    return t < o.t;
  }
};

// W. R. Franklin's crossing-number test: is (x, y) inside the polygon?
bool inside(int n, const float *px, const float *py, float x, float y) {
  bool c = false;
  for (int i = 0, j = n - 1; i < n; j = i++)
    if ((py[i] > y) != (py[j] > y) && x < (px[j] - px[i]) * (y - py[i]) / (py[j] - py[i]) + px[i])
      c = !c;
  return c;
}

// Triangles of one face as tinyobjloader's LoadObj(triangulate = true)
// makes them (common/3rdParty/tiny_obj_loader.h:1385-1590): a triangle
// stays as it is; a polygon is projected on the two axes the normal of its
// first non-degenerate corner favours and cut by ear clipping from its first
// corner on (a convex polygon becomes the fan of its first corner).  A
// polygon that runs out of ears keeps the triangles cut so far.
std::vector<std::array<Corner, 3>> triangulate(const std::vector<Corner> &poly,
                                               const std::vector<float> &pos) {
  std::vector<std::array<Corner, 3>> out;
  const size_t n = poly.size();
  if (n == 3) {
    out.push_back({poly[0], poly[1], poly[2]});
    return out;
  }
  auto p = [&](const Corner &c, int axis) { return pos[3 * (size_t)c.v + (size_t)axis]; };
  int ax = 1, ay = 2;  // projection axes
  for (size_t k = 0; k < n; ++k) {
    const Corner &a = poly[k], &b = poly[(k + 1) % n], &c = poly[(k + 2) % n];
    float e0[3], e1[3];
    for (int i = 0; i < 3; ++i) e0[i] = p(b, i) - p(a, i), e1[i] = p(c, i) - p(b, i);
    const float cx = std::fabs(e0[1] * e1[2] - e0[2] * e1[1]);
    const float cy = std::fabs(e0[2] * e1[0] - e0[0] * e1[2]);
    const float cz = std::fabs(e0[0] * e1[1] - e0[1] * e1[0]);
    const float eps = std::numeric_limits<float>::epsilon();
    if (cx > eps || cy > eps || cz > eps) {
      if (!(cx > cy && cx > cz)) ax = 0, ay = cz > cx && cz > cy ? 1 : 2;
      break;
    }
  }
  float area = 0.f;  // signed area of the projection
  for (size_t k = 0; k < n; ++k) {
    const Corner &a = poly[k], &b = poly[(k + 1) % n];
    area += (p(a, ax) * p(b, ay) - p(a, ay) * p(b, ax)) * 0.5f;
  }
  std::vector<Corner> rest = poly;
  size_t guess = 0, budget = n, last = n;
  while (rest.size() > 3 && budget > 0) {
    const size_t m = rest.size();
    if (guess >= m) guess -= m;
    if (last != m) {
      last = m;
      budget = m;
    } else {
      --budget;
    }
    Corner tri[3];
    float x[3], y[3];
    for (int k = 0; k < 3; ++k) {
      tri[k] = rest[(guess + (size_t)k) % m];
      x[k] = p(tri[k], ax), y[k] = p(tri[k], ay);
    }
    const float turn = (x[1] - x[0]) * (y[2] - y[1]) - (y[1] - y[0]) * (x[2] - x[1]);
    bool ear = !(turn * area < 0.f);  // a reflex corner is no ear
    for (size_t o = 3; ear && o < m; ++o) {
      const Corner &q = rest[(guess + o) % m];
      if (inside(3, x, y, p(q, ax), p(q, ay))) ear = false;
    }
    if (!ear) {
      ++guess;
      continue;
    }
    out.push_back({tri[0], tri[1], tri[2]});
    rest.erase(rest.begin() + (long)((guess + 1) % m));
  }
  if (rest.size() == 3) out.push_back({rest[0], rest[1], rest[2]});
  return out;
}

ObjMesh read_obj(const std::string &path) {
  std::ifstream in(path);
  require((bool)in, NWPRT_ERR_IO, "obj: cannot open " + path);
  const std::string dir = [&] {
    const size_t s = path.find_last_of('/');
    return s == std::string::npos ? std::string() : path.substr(0, s + 1);
  }();
  std::vector<float> pos, nrm, tex;
  std::map<std::string, int> names;
  std::vector<nwprt_material> mats;
  std::vector<std::string> maps;
  struct Face {
    Corner c[3];
    int shape, mat;
  };
  std::vector<Face> faces;
  int cur_mat = -1, shape = 0;
  bool shape_has_faces = false;
  auto resolve = [](long v, size_t count) -> int {
    const long n = (long)count;
    const long i = v > 0 ? v - 1 : n + v;
    require(v != 0 && i >= 0 && i < n, NWPRT_ERR_FORMAT, "obj: index out of range");
    return (int)i;
  };
  size_t line_no = 0;
  for (std::string line; std::getline(in, line);) {
    ++line_no;
    const auto t = split_ws(line);
    if (t.empty() || t[0][0] == '#') continue;
    try {
      if (t[0] == "v" && t.size() >= 4) {
        for (int k = 1; k <= 3; ++k) pos.push_back(std::stof(t[(size_t)k]));
      } else if (t[0] == "vn" && t.size() >= 4) {
        for (int k = 1; k <= 3; ++k) nrm.push_back(std::stof(t[(size_t)k]));
      } else if (t[0] == "vt" && t.size() >= 3) {
        for (int k = 1; k <= 2; ++k) tex.push_back(std::stof(t[(size_t)k]));
      } else if (t[0] == "mtllib" && t.size() > 1) {
        read_mtl(dir + t[1], dir, names, mats, maps);
      } else if (t[0] == "usemtl" && t.size() > 1) {
        auto it = names.find(t[1]);
        cur_mat = it == names.end() ? -1 : it->second;
      } else if (t[0] == "o" || t[0] == "g") {
        // tinyobj starts a new shape at an object or group line.
        if (shape_has_faces) ++shape;
        shape_has_faces = false;
      } else if (t[0] == "f" && t.size() >= 4) {
        std::vector<Corner> poly;
        for (size_t c = 1; c < t.size(); ++c) {
          const std::string &s = t[c];
          const size_t a = s.find('/');
          Corner k{resolve(std::stol(s.substr(0, a)), pos.size() / 3), -1, -1};
          if (a != std::string::npos) {
            const size_t b = s.find('/', a + 1);
            const std::string ts = s.substr(a + 1, b == std::string::npos ? std::string::npos : b - a - 1);
            if (!ts.empty()) k.t = resolve(std::stol(ts), tex.size() / 2);
            if (b != std::string::npos && b + 1 < s.size())
              k.n = resolve(std::stol(s.substr(b + 1)), nrm.size() / 3);
          }
          poly.push_back(k);
        }
        for (const auto &tri : triangulate(poly, pos))
          faces.push_back(Face{{tri[0], tri[1], tri[2]}, shape, cur_mat});
        shape_has_faces = true;
      }
    } catch (const Error &) {
      throw;
    } catch (const std::exception &) {
      fail(NWPRT_ERR_FORMAT, "obj: bad line " + std::to_string(line_no) + " of " + path);
    }
  }
  require(!faces.empty(), NWPRT_ERR_FORMAT, "obj: no faces in " + path);
  // Faces without a material take cookTorrance's defaults (kd = ks = 0.5,
  // roughness 1), as one more material.
  int default_mat = -1;
  for (const Face &f : faces)
    if (f.mat < 0 && default_mat < 0) {
      default_mat = (int)mats.size();
      mats.push_back(cook_torrance(0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 1.f));
      maps.resize(mats.size() * 3);
    }
  maps.resize(mats.size() * 3);
  // Meshes in the reference's order: per shape, one mesh per material in
  // ascending material id (faces without one first), faces in file order;
  // each mesh numbers its vertices by first use of a distinct (position,
  // normal, texcoord) corner, and the scene's vertices are the meshes'
  // vertices one after the other (Model.cpp:73-146, OptixBackend.cpp:700-752).
  ObjMesh m;
  const bool any_n = [&] {
    for (const Face &f : faces)
      for (const Corner &c : f.c)
        if (c.n >= 0) return true;
    return false;
  }();
  const bool any_t = [&] {
    for (const Face &f : faces)
      for (const Corner &c : f.c)
        if (c.t >= 0) return true;
    return false;
  }();
  std::map<std::pair<int, int>, std::vector<size_t>> groups;  // (shape, material) -> faces
  for (size_t i = 0; i < faces.size(); ++i) groups[{faces[i].shape, faces[i].mat}].push_back(i);
  for (const auto &g : groups) {  // shape, then material id (-1 first)
    std::map<Corner, int32_t> known;
    const size_t first = m.vertices.size() / 3;
    std::vector<float> gn, gt;  // this mesh's normals / texcoords, back-filled
    for (size_t fi : g.second) {
      for (const Corner &c : faces[fi].c) {
        auto it = known.find(c);
        if (it == known.end()) {
          const int32_t id = (int32_t)(m.vertices.size() / 3);
          for (int k = 0; k < 3; ++k) m.vertices.push_back(pos[3 * (size_t)c.v + k]);
          const size_t local = (size_t)id - first + 1;
          // addVertex: a corner with a normal / texcoord fills every vertex
          // of the mesh so far that has none; the arrays then match the
          // vertex count (zeros for trailing vertices without one).
          if (c.n >= 0)
            while (gn.size() < 3 * local)
              for (int k = 0; k < 3; ++k) gn.push_back(nrm[3 * (size_t)c.n + k]);
          if (c.t >= 0)
            while (gt.size() < 2 * local)
              for (int k = 0; k < 2; ++k) gt.push_back(tex[2 * (size_t)c.t + k]);
          if (!gt.empty()) gt.resize(2 * local, 0.f);
          if (!gn.empty()) gn.resize(3 * local, 0.f);
          it = known.emplace(c, id).first;
        }
        m.indices.push_back(it->second);
      }
      m.material_ids.push_back(g.first.second < 0 ? default_mat : g.first.second);
    }
    const size_t count = m.vertices.size() / 3 - first;
    gn.resize(3 * count, 0.f);
    gt.resize(2 * count, 0.f);
    if (any_n) m.normals.insert(m.normals.end(), gn.begin(), gn.end());
    if (any_t) m.texcoords.insert(m.texcoords.end(), gt.begin(), gt.end());
  }
  m.materials = mats;
  m.maps = maps;
  m.has_normals = any_n;
  m.has_texcoords = any_t;
  return m;
}

}  // namespace

// ------------------------------------------------------------------ PLY --
// Stanford PLY reader for the same mesh arrays, following what loadPly
// (Renderer/Model.cpp:152-220, tinyply) keeps: vertex x y z, optional
// nx ny nz, and the face element's vertex_indices.  ASCII and binary
// little-endian files; any scalar property type; other elements and
// properties are skipped.  Faces with more than three corners are fanned
// (loadPly reads triangles only).  A PLY file has no material: the mesh gets
// cookTorranceMaterial's defaults, as loadPly's caller gives it
// (Model.cpp:231-235).
namespace {

struct PlyProp {
  std::string name, type, count_type;  // count_type non-empty: a list
};
struct PlyElem {
  std::string name;
  int64_t count = 0;
  std::vector<PlyProp> props;
};

int ply_type_size(const std::string &t) {
  if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
  if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
  if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" ||
      t == "float32")
    return 4;
  if (t == "double" || t == "float64") return 8;
  fail(NWPRT_ERR_FORMAT, "ply: unknown property type " + t);
}

double ply_binary_value(const std::string &t, const unsigned char *p) {
  const int n = ply_type_size(t);
  if (t == "float" || t == "float32") { float v; std::memcpy(&v, p, 4); return v; }
  if (t == "double" || t == "float64") { double v; std::memcpy(&v, p, 8); return v; }
  const bool sgn = t == "char" || t == "int8" || t == "short" || t == "int16" || t == "int" ||
                   t == "int32";
  if (n == 1) return sgn ? (double)(int8_t)p[0] : (double)p[0];
  if (n == 2) {
    uint16_t v;
    std::memcpy(&v, p, 2);
    return sgn ? (double)(int16_t)v : (double)v;
  }
  uint32_t v;
  std::memcpy(&v, p, 4);
  return sgn ? (double)(int32_t)v : (double)v;
}

ObjMesh read_ply(const std::string &path) {
  std::ifstream in(path, std::ios::binary);
  require((bool)in, NWPRT_ERR_IO, "ply: cannot open " + path);
  std::string line;
  require((bool)std::getline(in, line) && line.rfind("ply", 0) == 0, NWPRT_ERR_FORMAT,
          "ply: missing magic in " + path);
  bool binary = false;
  std::vector<PlyElem> elems;
  for (;;) {
    require((bool)std::getline(in, line), NWPRT_ERR_FORMAT, "ply: header not terminated");
    if (!line.empty() && line.back() == '\r') line.pop_back();
    const auto t = split_ws(line);
    if (t.empty() || t[0] == "comment" || t[0] == "obj_info") continue;
    if (t[0] == "end_header") break;
    if (t[0] == "format") {
      require(t.size() >= 2, NWPRT_ERR_FORMAT, "ply: bad format line");
      if (t[1] == "binary_little_endian") binary = true;
      else require(t[1] == "ascii", NWPRT_ERR_FORMAT, "ply: unsupported format " + t[1]);
    } else if (t[0] == "element" && t.size() >= 3) {
      PlyElem e;
      e.name = t[1];
      e.count = std::stoll(t[2]);
      require(e.count >= 0, NWPRT_ERR_FORMAT, "ply: negative element count");
      elems.push_back(e);
    } else if (t[0] == "property" && !elems.empty()) {
      PlyProp pr;
      if (t.size() >= 5 && t[1] == "list") {
        pr.count_type = t[2];
        pr.type = t[3];
        pr.name = t[4];
      } else {
        require(t.size() >= 3, NWPRT_ERR_FORMAT, "ply: bad property line");
        pr.type = t[1];
        pr.name = t[2];
      }
      ply_type_size(pr.type);
      if (!pr.count_type.empty()) ply_type_size(pr.count_type);
      elems.back().props.push_back(pr);
    }
  }
  std::vector<double> pos, nrm;
  std::vector<std::vector<int64_t>> faces;
  bool has_normal = false;
  for (const PlyElem &e : elems) {
    int ix = -1, iy = -1, iz = -1, inx = -1, iny = -1, inz = -1, ifc = -1;
    for (size_t k = 0; k < e.props.size(); ++k) {
      const std::string &n = e.props[k].name;
      if (n == "x") ix = (int)k;
      if (n == "y") iy = (int)k;
      if (n == "z") iz = (int)k;
      if (n == "nx") inx = (int)k;
      if (n == "ny") iny = (int)k;
      if (n == "nz") inz = (int)k;
      if ((n == "vertex_indices" || n == "vertex_index") && !e.props[k].count_type.empty())
        ifc = (int)k;
    }
    const bool is_vertex = e.name == "vertex", is_face = e.name == "face";
    if (is_vertex) {
      require(ix >= 0 && iy >= 0 && iz >= 0, NWPRT_ERR_FORMAT, "ply: vertex needs x y z");
      has_normal = inx >= 0 && iny >= 0 && inz >= 0;
    }
    if (is_face) require(ifc >= 0, NWPRT_ERR_FORMAT, "ply: face needs vertex_indices");
    std::vector<double> vals(e.props.size());
    std::vector<int64_t> list;
    for (int64_t r = 0; r < e.count; ++r) {
      std::vector<std::string> tok;
      size_t tk = 0;
      if (!binary) {
        require((bool)std::getline(in, line), NWPRT_ERR_FORMAT, "ply: truncated ascii data");
        tok = split_ws(line);
      }
      auto next = [&](const std::string &type) -> double {
        if (binary) {
          unsigned char b[8];
          const int n = ply_type_size(type);
          in.read((char *)b, n);
          require((bool)in, NWPRT_ERR_FORMAT, "ply: truncated binary data");
          return ply_binary_value(type, b);
        }
        require(tk < tok.size(), NWPRT_ERR_FORMAT, "ply: short ascii record");
        return std::stod(tok[tk++]);
      };
      for (size_t k = 0; k < e.props.size(); ++k) {
        const PlyProp &pr = e.props[k];
        if (pr.count_type.empty()) {
          vals[k] = next(pr.type);
        } else {
          const double c = next(pr.count_type);
          require(c >= 0 && c < 1e6, NWPRT_ERR_FORMAT, "ply: bad list length");
          list.clear();
          for (int64_t j = 0; j < (int64_t)c; ++j) list.push_back((int64_t)next(pr.type));
          if (is_face && (int)k == ifc) faces.push_back(list);
        }
      }
      if (is_vertex) {
        pos.insert(pos.end(), {vals[ix], vals[iy], vals[iz]});
        if (has_normal) nrm.insert(nrm.end(), {vals[inx], vals[iny], vals[inz]});
      }
    }
  }
  const int64_t nv = (int64_t)(pos.size() / 3);
  ObjMesh m;
  for (double v : pos) m.vertices.push_back((float)v);
  for (double v : nrm) m.normals.push_back((float)v);
  m.has_normals = has_normal;
  for (const auto &f : faces) {
    require(f.size() >= 3, NWPRT_ERR_FORMAT, "ply: face with fewer than 3 corners");
    for (int64_t c : f)
      require(c >= 0 && c < nv, NWPRT_ERR_FORMAT, "ply: vertex index out of range");
    for (size_t c = 1; c + 1 < f.size(); ++c) {
      m.indices.push_back((int32_t)f[0]);
      m.indices.push_back((int32_t)f[c]);
      m.indices.push_back((int32_t)f[c + 1]);
      m.material_ids.push_back(0);
    }
  }
  require(!m.indices.empty(), NWPRT_ERR_FORMAT, "ply: no faces in " + path);
  m.materials.push_back(cook_torrance(0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 1.f));
  return m;
}

}  // namespace

extern "C" nwprt_status nwprt_ply_read(const char *path, float **vertices, int64_t *n_vertices,
                                       float **normals, int32_t **indices, int64_t *n_triangles,
                                       int32_t **material_ids, nwprt_material **materials,
                                       int32_t *n_materials) {
  return host_guarded([&] {
    require(path && vertices && n_vertices && normals && indices && n_triangles &&
                material_ids && materials && n_materials,
            NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    const ObjMesh m = read_ply(path);
    *vertices = copy_out(m.vertices);
    *normals = m.has_normals ? copy_out(m.normals) : nullptr;
    *indices = copy_out(m.indices);
    *material_ids = copy_out(m.material_ids);
    *materials = copy_out(m.materials);
    *n_vertices = (int64_t)(m.vertices.size() / 3);
    *n_triangles = (int64_t)(m.indices.size() / 3);
    *n_materials = (int32_t)m.materials.size();
  });
}

extern "C" nwprt_status nwprt_obj_read(const char *path, float **vertices, int64_t *n_vertices,
                                       float **normals, int32_t **indices, int64_t *n_triangles,
                                       int32_t **material_ids, nwprt_material **materials,
                                       int32_t *n_materials) {
  return host_guarded([&] {
    require(path && vertices && n_vertices && normals && indices && n_triangles &&
                material_ids && materials && n_materials,
            NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    const ObjMesh m = read_obj(path);
    *vertices = copy_out(m.vertices);
    *normals = m.has_normals ? copy_out(m.normals) : nullptr;
    *indices = copy_out(m.indices);
    *material_ids = copy_out(m.material_ids);
    *materials = copy_out(m.materials);
    *n_vertices = (int64_t)(m.vertices.size() / 3);
    *n_triangles = (int64_t)(m.indices.size() / 3);
    *n_materials = (int32_t)m.materials.size();
  });
}

extern "C" nwprt_status nwprt_image_read_rgba8(const char *path, uint8_t **rgba, int32_t *width,
                                               int32_t *height) {
  return host_guarded([&] {
    require(path && rgba && width && height, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    const Image img = read_image_rgba8(path);
    *rgba = copy_out(img.rgba);
    *width = img.width;
    *height = img.height;
  });
}

extern "C" nwprt_status nwprt_obj_read_textured(const char *path, float **vertices,
                                                int64_t *n_vertices, float **normals,
                                                float **texcoords, int32_t **indices,
                                                int64_t *n_triangles, int32_t **material_ids,
                                                nwprt_material **materials, int32_t *n_materials,
                                                char **map_paths, int64_t *map_bytes) {
  return host_guarded([&] {
    require(path && vertices && n_vertices && normals && texcoords && indices && n_triangles &&
                material_ids && materials && n_materials && map_paths && map_bytes,
            NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    const ObjMesh m = read_obj(path);
    std::vector<char> block;
    for (const std::string &f : m.maps) {
      block.insert(block.end(), f.begin(), f.end());
      block.push_back('\0');
    }
    *vertices = copy_out(m.vertices);
    *normals = m.has_normals ? copy_out(m.normals) : nullptr;
    *texcoords = m.has_texcoords ? copy_out(m.texcoords) : nullptr;
    *indices = copy_out(m.indices);
    *material_ids = copy_out(m.material_ids);
    *materials = copy_out(m.materials);
    *map_paths = copy_out(block);
    *map_bytes = (int64_t)block.size();
    *n_vertices = (int64_t)(m.vertices.size() / 3);
    *n_triangles = (int64_t)(m.indices.size() / 3);
    *n_materials = (int32_t)m.materials.size();
  });
}

extern "C" nwprt_status nwprt_material_defaults(int32_t type, nwprt_material *out) {
  return host_guarded([&] {
    require(out && type >= 0 && type <= 4, NWPRT_ERR_INVALID_ARGUMENT,
            "material_defaults: null output or type outside 0..4");
    *out = material_defaults(type);
  });
}
