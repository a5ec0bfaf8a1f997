// Internal declarations shared by the host sources and kernels of libnwprt.
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nwprt.h"

namespace nwprt {

// ---------------------------------------------------------------- errors --

// Carries an nwprt_status across the C++ layers; converted to a status code
// and a thread-local message at the C ABI.
struct Error : std::runtime_error {
  nwprt_status status;
  Error(nwprt_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] void fail(nwprt_status s, const std::string &msg);
void require(bool cond, nwprt_status s, const std::string &msg);
void set_last_error(const std::string &msg);
const char *last_error_cstr();

// ------------------------------------------------------------------ fp16 --

// IEEE binary16 / bfloat16 conversions, round to nearest even.
uint16_t f32_to_f16_bits(float f);
float f16_bits_to_f32(uint16_t h);
uint16_t f32_to_bf16_bits(float f);

// ------------------------------------------------------------- hashgrid --

void hashgrid_levels(const nwprt_hashgrid_desc &d, nwprt_hashgrid_levels &out);
void validate_scene_desc(const nwprt_scene_desc &d);
size_t scene_hash_entries(const nwprt_scene_desc &d);
size_t scene_layer_weight_count(const nwprt_scene_desc &d);

// ------------------------------------------------------------- formats --

struct NpyArray {
  std::vector<int64_t> shape;
  bool fortran_order = false;
  char kind = 'f';       // 'f', 'i', 'u', 'b'
  int itemsize = 4;
  std::vector<uint8_t> raw;  // little-endian payload
  size_t count() const;
  std::vector<float> as_f32() const;     // values in stored (file) order
  std::vector<int32_t> as_i32() const;
};
NpyArray npy_read(const std::string &path);
void npy_write(const std::string &path, const void *data, char kind,
               int itemsize, const std::vector<int64_t> &shape);

struct HdrImage {
  int width = 0, height = 0;
  std::vector<float> rgb;  // rows top first (file order)
};
HdrImage hdr_read(const std::string &path);
void hdr_write(const std::string &path, const float *data, int width,
               int height, int channels);
void cross_to_faces(const float *cross, int width, int height, float *faces);

// An 8-bit RGBA picture with row 0 at the bottom (image.cpp).
struct Image {
  int width = 0, height = 0;
  std::vector<uint8_t> rgba;
};
Image read_image_rgba8(const std::string &path);

// Minimal JSON value for config_comp.json.
struct Json {
  enum Type { Null, Bool, Number, String, Array, Object } type = Null;
  bool b = false;
  double num = 0;
  std::string str;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;
  const Json *get(const std::string &key) const;
  double number_or(const std::string &key, double dflt) const;
  std::string string_or(const std::string &key, const std::string &dflt) const;
};
Json json_parse(const std::string &text);

// ------------------------------------------------------------- synthetic --

void synth_scene(const nwprt_scene_desc &d, uint64_t seed, float hash_std,
                 float feature_std, void *hash_table, float *wfeat,
                 float *proj, float *layers);
void synth_gbuffer(uint64_t seed, int w, int h, float background_fraction,
                   const nwprt_gbuffer &out);
void synth_cubemap(uint64_t seed, int n, float *faces);

// Loads a reference weight folder into host arrays (see model_folder.cpp).
struct HostScene {
  nwprt_scene_desc desc{};
  std::vector<uint8_t> hash_table;  // entries x F in desc.hash_dtype
  std::vector<float> wavelet_features;
  std::vector<float> projection;    // empty if !has_projection
  std::vector<float> layers;
};
HostScene load_model_folder(const std::string &dir);

}  // namespace nwprt
