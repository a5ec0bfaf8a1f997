// Host-only helpers: error plumbing, FP16/BF16 rounding, hash-grid level
// table and scene-shape validation.  Nothing here touches the GPU.
#include <cmath>
#include <cstdio>

#include "nwprt_internal.hpp"

namespace nwprt {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }
const char *last_error_cstr() { return g_last_error.c_str(); }

void fail(nwprt_status s, const std::string &msg) { throw Error(s, msg); }

void require(bool cond, nwprt_status s, const std::string &msg) {
  if (!cond) fail(s, msg);
}

// ------------------------------------------------------------------ fp16 --

uint16_t f32_to_f16_bits(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t absx = x & 0x7fffffffu;
  if (absx >= 0x7f800000u) {  // inf / nan
    return (uint16_t)(sign | 0x7c00u | (absx > 0x7f800000u ? 0x200u : 0u));
  }
  if (absx >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds to inf
  if (absx < 0x38800000u) {  // result is subnormal or zero in binary16
    if (absx < 0x33000000u) return (uint16_t)sign;  // < 2^-25: rounds to 0
    const uint32_t e = absx >> 23;
    const uint32_t m = (absx & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126u - e;  // 14..24 for the range above
    uint32_t q = m >> shift;
    const uint32_t rem = m & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  uint32_t r = absx - 0x38000000u;  // rebias exponent 127 -> 15
  const uint32_t rem = r & 0x1fffu;
  r >>= 13;
  if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) r++;
  return (uint16_t)(sign | r);
}

float f16_bits_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu;
  uint32_t m = h & 0x3ffu;
  uint32_t x;
  if (e == 0) {
    if (m == 0) {
      x = sign;
    } else {  // subnormal: normalise
      e = 113;  // 127 - 15 + 1
      while ((m & 0x400u) == 0) {
        m <<= 1;
        e--;
      }
      m &= 0x3ffu;
      x = sign | (e << 23) | (m << 13);
    }
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

uint16_t f32_to_bf16_bits(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  const uint32_t lsb = (x >> 16) & 1u;
  x += 0x7fffu + lsb;
  return (uint16_t)(x >> 16);
}

// ------------------------------------------------------------- hashgrid --

// Level table of the multiresolution hash grid, evaluated in FP32 on the
// host exactly once (tiny-cuda-nn's conventions as used through
// Renderer/tinycudann/prt_network.h:45 and :86-91):
//   scale_l = exp2(l * log2(b)) * N_min - 1,   res_l = ceil(scale_l) + 1,
//   dense when res_l^3 <= T, size_l = min(next_multiple(res_l^3, 8), T).
void hashgrid_levels(const nwprt_hashgrid_desc &d, nwprt_hashgrid_levels &out) {
  require(d.n_levels >= 1 && d.n_levels <= NWPRT_MAX_LEVELS,
          NWPRT_ERR_INVALID_ARGUMENT, "hash grid: n_levels out of range");
  require(d.n_features_per_level >= 1 && d.n_features_per_level <= 8,
          NWPRT_ERR_INVALID_ARGUMENT, "hash grid: n_features_per_level out of range");
  require(d.log2_hashmap_size >= 3 && d.log2_hashmap_size <= 26,
          NWPRT_ERR_INVALID_ARGUMENT, "hash grid: log2_hashmap_size out of range");
  require(d.base_resolution >= 1 && d.per_level_scale >= 1.0f,
          NWPRT_ERR_INVALID_ARGUMENT, "hash grid: bad base resolution / scale");
  std::memset(&out, 0, sizeof(out));
  out.n_levels = d.n_levels;
  const uint64_t T = 1ull << d.log2_hashmap_size;
  const float log2_b = std::log2(d.per_level_scale);  // float overload
  uint64_t offset = 0;
  for (int l = 0; l < d.n_levels; ++l) {
    const float scale =
        std::exp2((float)l * log2_b) * (float)d.base_resolution - 1.0f;
    const uint32_t res = (uint32_t)std::ceil(scale) + 1u;
    const uint64_t cube = (uint64_t)res * res * res;
    const bool dense = cube <= T;
    uint64_t size = dense ? ((cube + 7) / 8) * 8 : T;
    if (size > T) size = T;
    out.scale[l] = scale;
    out.resolution[l] = res;
    out.dense[l] = dense ? 1u : 0u;
    out.size[l] = (uint32_t)size;
    out.offset[l] = (uint32_t)offset;
    offset += size;
    require(offset < (1ull << 31), NWPRT_ERR_INVALID_ARGUMENT,
            "hash grid: total entries exceed 2^31");
  }
  out.offset[d.n_levels] = (uint32_t)offset;
}

void validate_scene_desc(const nwprt_scene_desc &d) {
  nwprt_hashgrid_levels lv;
  hashgrid_levels(d.grid, lv);
  require(d.hash_dtype == NWPRT_DTYPE_F32 || d.hash_dtype == NWPRT_DTYPE_F16,
          NWPRT_ERR_INVALID_ARGUMENT, "scene: unknown hash_dtype");
  require(d.face_size >= 2 && d.face_size <= NWPRT_MAX_FACE_SIZE &&
              (d.face_size & (d.face_size - 1)) == 0,
          NWPRT_ERR_INVALID_ARGUMENT,
          "scene: face_size must be a power of two in [2, 128]");
  require(d.feature_dim == d.grid.n_levels * d.grid.n_features_per_level,
          NWPRT_ERR_INVALID_ARGUMENT,
          "scene: feature_dim must equal n_levels * n_features_per_level");
  require(d.feature_dim == 32 && d.hidden_width == 64 && d.n_hidden_layers == 3,
          NWPRT_ERR_UNSUPPORTED,
          "scene: this build implements feature_dim 32, hidden_width 64, "
          "3 hidden layers");
  require(d.has_projection == 0 || d.has_projection == 1,
          NWPRT_ERR_INVALID_ARGUMENT, "scene: has_projection must be 0/1");
  require(d.output_softplus == 0 || d.output_softplus == 1,
          NWPRT_ERR_INVALID_ARGUMENT, "scene: output_softplus must be 0/1");
  require(d.aux_dims == 0 || d.aux_dims == 13, NWPRT_ERR_INVALID_ARGUMENT,
          "scene: aux_dims must be 0 or 13");
}

size_t scene_hash_entries(const nwprt_scene_desc &d) {
  nwprt_hashgrid_levels lv;
  hashgrid_levels(d.grid, lv);
  return (size_t)lv.offset[lv.n_levels];
}

size_t scene_layer_weight_count(const nwprt_scene_desc &d) {
  const size_t H = (size_t)d.hidden_width, F = (size_t)d.feature_dim;
  return H * F + (size_t)(d.n_hidden_layers - 1) * H * H + 3 * H +
         (d.aux_dims ? H * 16 : 0);
}

}  // namespace nwprt
