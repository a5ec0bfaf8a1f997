// C ABI of the host-only utilities (no GPU use): level table, parameter
// counts, seeded inputs and file formats.
#include <cstdlib>

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
