// Launch interfaces of the gfx950 kernels (implemented in the *.hip files).
// Host code fills these argument structs; every launcher checks its shapes
// before launching and throws nwprt::Error on a HIP failure.
#pragma once

#include <hip/hip_runtime.h>

#include "nwprt_internal.hpp"

namespace nwprt {

void hip_check(hipError_t e, const char *what);

// Element type of per-pixel features and MFMA operands.
enum class Elem : int { F32 = 0, F16 = 1, BF16 = 2 };

// ---------------------------------------------------------------- envmap --

// Non-standard 2-D Haar of `n_maps` cubemaps (faces [6][N][N][3] FP32 each,
// maps `map_stride` floats apart) into coefficients of the same layout.
// One workgroup per (map, face, channel); scaling coefficients in FP64.
void launch_haar(const float *faces, size_t map_stride, float *coef, int n,
                 int n_maps, hipStream_t s);

// Area-weighted top-K of one map's coefficients: indices int32[K] in
// descending (score, then ascending index) order and strengths float[K][3].
// `scores` is a uint32[6N^2] work buffer.
void launch_topk(const float *coef, int n, int k, uint32_t *scores,
                 int32_t *indices, float *strengths, hipStream_t s);

// Per-relight preparation of the selected wavelets:
//  mk   (may be null): [K][4][64][8] MFMA A-operand tiles of W1 * diag(wfeat[:,c_k])
//  wsel (may be null): [K][32] FP32 selected wavelet features.
void launch_prepare_selection(const float *w1, const float *wfeat, int n_coef,
                              const int32_t *indices, int k, Elem mk_elem,
                              uint16_t *mk, float *wsel, hipStream_t s);

// ---------------------------------------------------------------- encode --

struct EncodeArgs {
  int n_pixels;
  const float *position;      // float4 per pixel
  const uint8_t *env_flag;    // may be null: all pixels encoded
  const void *hash;           // entries x F
  int hash_f16;               // 1: FP16 entries, 0: FP32
  int n_features;             // F per level (1, 2, 4 or 8); L = 32 / F
  nwprt_hashgrid_levels lv;
  const float *projection;    // [32][32] or null (identity)
  void *features;             // [n][32] in out_elem
  Elem out_elem;
  // Aux G-buffer encodings (null aux: none): 13 values per pixel from the
  // G-buffer plus 3 padding values, written as [n][16] in out_elem.
  const float *normal, *reflect, *roughness, *kd, *ks;
  void *aux;
  float aux_pad;
  // debug tap (null when unused)
  uint32_t *dbg_indices;      // [n][L][8]
  float *dbg_features;        // [n][32] unprojected FP32
};
void launch_encode(const EncodeArgs &a, hipStream_t s);

// --------------------------------------------------------------- relight --

struct RelightArgs {
  int n_pixels;
  int k;
  const void *features;       // [P][32], Elem per kernel
  const uint16_t *mk;         // MFMA: [K][4][64][8]
  const float *wsel;          // plain: [K][32]
  const float *light;         // [K][3]
  const uint16_t *w_tiles;    // MFMA: W2 | W3 | W4 variants, [4][2][64][8] each
  const float *w_f32;         // plain: W1 [64][32] | W2 | W3 [64][64] | W4 [3][64]
  // Aux inputs of the second layer (null: none): per-pixel [P][16] encodings
  // (Elem per kernel) and the second layer's aux columns, as MFMA A tiles
  // [4][64][4] (fused) or FP32 [64][16] (plain).
  const void *aux;
  const uint16_t *waux_tiles;
  const float *waux_f32;
  const uint8_t *env_flag;
  const float *direct;        // float4
  float *indirect;            // float4, may be null
  float *with_direct;         // float4, may be null
  uint32_t *rgba8;            // may be null
  int tonemap_source;         // nwprt_tonemap_source
  int softplus;
  // transport tap: T for pixels [dump_begin, dump_begin + dump_count)
  float *dump;
  int dump_begin, dump_count;
};
// Tile-shape variants of the fused kernel (see relight_kernels.hip).
int relight_mfma_variant_count();
const char *relight_mfma_variant_name(int variant);
void launch_relight_mfma(const RelightArgs &a, Elem elem, int variant, hipStream_t s);
void launch_relight_plain(const RelightArgs &a, hipStream_t s);
void launch_tonemap(const float *src, uint32_t *dst, int n_pixels, hipStream_t s);

// Host-side MFMA weight tiling of layers 2..4 and of the second layer's aux
// columns (see relight kernels).
void build_weight_tiles(const float *w_f32, Elem elem, std::vector<uint16_t> &out);
void build_aux_tiles(const float *waux, Elem elem, std::vector<uint16_t> &out);

}  // namespace nwprt
