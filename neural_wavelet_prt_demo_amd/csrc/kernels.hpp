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

// Work buffer of the Haar transform (n_maps maps) and of the top-K (one map)
// at face size n: envmap_work_bytes(n, n_maps) bytes.  The two may share one
// buffer when they run in stream order.
constexpr int kScoreKeys = 2048;  // keys per score workgroup
size_t envmap_work_bytes(int n, int n_maps);

// Non-standard 2-D Haar of `n_maps` cubemaps (faces [6][N][N][3] FP32 each,
// maps `map_stride` floats apart) into coefficients of the same layout:
// tiles of 32 x 32 texels per workgroup, then the levels above the tiles;
// scaling coefficients in FP64.
void launch_haar(const float *faces, size_t map_stride, float *coef, int n,
                 int n_maps, void *work, hipStream_t s);

// Area-weighted top-K of one map's coefficients: indices int32[K] in
// descending (score, then ascending index) order and strengths float[K][3].
void launch_topk(const float *coef, int n, int k, void *work,
                 int32_t *indices, float *strengths, hipStream_t s);

// Yaw rotation (radians, about +y) of a cubemap [6][N][N][3] into `dst`
// (nearest source texel to each rotated texel centre).
void launch_rotate_cubemap(const float *src, float *dst, int n, double yaw, hipStream_t s);

// Per-relight preparation of the selected wavelets:
//  mk   (may be null): [K][4][64][8] 16x16x32 MFMA A-operand tiles of
//                      W1 * diag(wfeat[:,c_k])
//  wsel (may be null): [K][32] FP32 selected wavelet features
//  mk32 (may be null): [K][2][2][64][8] BF16 32x32x16 A-operand tiles of the
//                      same product, for the FP8 form's first layer.
// planar: the MFMA tiles are for planar features -- contraction slot j
// carries raw feature planar_feature(j), with the [32][32] projection (if
// not null) folded in: W1 diag(wfeat[:,c_k]) P, columns permuted.
void launch_prepare_selection(const float *w1, const float *wfeat, int n_coef,
                              const int32_t *indices, int k, Elem mk_elem,
                              uint16_t *mk, float *wsel, hipStream_t s,
                              uint16_t *mk32 = nullptr, bool planar = false,
                              const float *projection = nullptr);

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
  // Planar form (16 levels x 2 features, no aux, no tap): `features` holds
  // the unprojected hash-grid features as 8 planes [8][n][4], plane q =
  // (level q f0, level q f1, level 15-q f0, level 15-q f1); the projection is
  // folded into the first-layer selection tiles instead.  Plane q is
  // computed by workgroups that land on XCD q, so each XCD's L2 serves two
  // levels' hash tables.
  int planar;
};
void launch_encode(const EncodeArgs &a, hipStream_t s);

// Raw hash-grid feature (2 level + feature) carried by slot `slot` in [0, 32)
// of the planar layout: slot / 4 is the plane, slot % 4 its value.
__host__ __device__ inline int planar_feature(int slot) {
  const int plane = slot >> 2, v = slot & 3;
  const int level = v < 2 ? plane : 15 - plane;
  return 2 * level + (v & 1);
}

// --------------------------------------------------------------- relight --

struct RelightArgs {
  int n_pixels;
  int k;
  const void *features;       // [P][32], Elem per kernel; planar: [8][P][4]
  int features_planar;        // MFMA forms: features in the planar layout
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
  // FP8 form: layers 2..4 as E4M3 A tiles [3][2][64][32] (W * 2^w8_exp per
  // layer), the calibrated activation maxima of layers 1..3 (device), and
  // with aux inputs the second layer's aux columns as BF16 32x32x16 A tiles
  // [2][64][8].
  const uint8_t *w8_tiles;
  int w8_exp[3];
  float *amax;
  const uint16_t *waux32_tiles;
  // Fused encode (one-call MFMA relights of 16 x 2 grids without aux
  // inputs): each wave computes its 64 pixels' hash-grid features from
  // `position` in the planar slot order instead of reading `features`.
  int fused_encode;
  const float *position;      // float4 per pixel
  const void *hash;           // entries x 2
  int hash_f16;
  nwprt_hashgrid_levels lv;
};
// Tile-shape variants of the fused kernel (see relight_kernels.hip).
int relight_mfma_variant_count();
const char *relight_mfma_variant_name(int variant);
void launch_relight_mfma(const RelightArgs &a, Elem elem, int variant, hipStream_t s);
void launch_relight_plain(const RelightArgs &a, hipStream_t s);
void launch_tonemap(const float *src, uint32_t *dst, int n_pixels, hipStream_t s);

// Guided a-trous denoiser (denoise_kernels.hip).
constexpr int kMaxDenoisePasses = 8;
struct DenoiseArgs {
  const float *color, *normal, *albedo;  // float4 [height][width]
  float *out;
  int width, height, iterations;
  float sigma_color, sigma_normal, sigma_albedo, blend;
};
// Device bytes of the work buffer launch_denoise needs for a w x h image.
size_t denoise_scratch_bytes(int w, int h);
void launch_denoise(const DenoiseArgs &a, void *scratch, hipStream_t s);

// G-buffer producer (gbuffer_kernels.hip, BVH built in mesh.cpp).
// The traversal stack holds at most one entry per internal level.
constexpr int kBvhMaxDepth = 24;
constexpr int kBvhLeafMax = 8;
// Two-child node, 64 bytes: both children's boxes sit in the parent so one
// visit tests both.  Child reference c >= 0: internal node c; c < 0: leaf
// ~c = first * 8 + (count - 1) over the reordered triangle array.
struct alignas(16) BvhNode {
  float lo0[3];
  int32_t c0;
  float hi0[3];
  int32_t c1;
  float lo1[3];
  int32_t pad0;
  float hi1[3];
  int32_t pad1;
};
// Triangle in leaf order: A, B - A, C - A and the original triangle index.
struct alignas(16) BvhTri {
  float v0[3];
  int32_t prim;
  float e1[3];
  int32_t pad0;
  float e2[3];
  int32_t pad1;
};
struct alignas(16) MaterialRec {
  float kd[3];
  float roughness;
  float ks[3];
  int32_t type;
  // DisneyBSDF parameters (Renderer/Material.h:45-55) and the index of
  // refraction roughPlastic and DisneyBSDF use.
  float specular_transmission, metallic, subsurface, specular;
  float specular_tint, anisotropic, sheen, sheen_tint;
  float clearcoat, clearcoat_gloss, eta, pad;
};
// A material texture as the G-buffer kernels sample it: RGBA8 texels, row 0
// at v = 0; rgba null: the field keeps the material's constant.  Fields
// (NWPRT_TEX_*): kd, ks (RGB), then roughness and the DisneyBSDF
// parameters and eta (first channel).
struct TexRef {
  const uint8_t *rgba;
  int32_t w, h;
};
static_assert(sizeof(TexRef) == 16, "TexRef layout");
constexpr int kTexFields = NWPRT_TEX_FIELDS;
struct CastArgs {
  const BvhNode *nodes;
  const BvhTri *tris;
  const int32_t *indices;       // [n_tri][3] original vertex indices
  const float *normals;         // [n_vertices][3], or null
  const int32_t *tri_material;  // [n_tri], or null (material 0)
  const MaterialRec *materials;
  float cam_pos[3], cam_dir[3], cam_h[3], cam_v[3];
  float cx, cy, fx, fy;
  float lower[3], box[3];
  int width, height, max_bounce;
  const float *env_faces;  // [6][env_n][env_n][3] or null
  int env_n;
  nwprt_gbuffer gb;  // null fields are skipped
  // Direct light (gb.direct and light_faces non-null): light_faces
  // [6][light_n][light_n][3] with its sampling tables env_tab (layout of
  // env_table_doubles), direct_samples estimates per pixel, RNG seed.
  const float *light_faces;
  const double *env_tab;
  int light_n, direct_samples;
  uint32_t seed;
  bool disney;  // some material is DisneyBSDF (selects the kernel instance)
  // Material textures: texcoords [n_vertices][2] and mat_tex
  // [n_materials][kTexFields], both null for an untextured mesh.
  const float *texcoords;
  const TexRef *mat_tex;
};
// Doubles of the environment sampling tables for face size n: face pmf[6],
// face cdf[7], then per face cdf_rows[n][n+1], pdf_rows[n][n],
// cdf_marginals[n+1], pdf_marginals[n].
size_t env_table_doubles(int n);
void launch_env_tables(const float *faces, int n, double *tab, hipStream_t s);
void launch_gbuffer_cast(const CastArgs &a, hipStream_t s);

// SH-PRT render (gbuffer_kernels.hip): per-vertex light response, colour
// per (vertex, material) pair of the triangle corners, per-pixel
// interpolation.
constexpr int kShMaxDegree = 8;
struct ShArgs {
  const BvhNode *nodes;
  const BvhTri *tris;
  const int32_t *indices;       // [n_tri][3]
  const float *vertices;        // [n_vertices][3]
  const float *normals;         // [n_vertices][3]
  const int32_t *tri_material;  // or null
  const MaterialRec *materials;
  int64_t n_tris, n_vertices;
  const float *transfer;        // [n_vertices][nc][nc][3], nc = (degree + 1)^2
  const float *light_sh;        // [nc][3]
  int degree, samples;
  uint32_t seed;
  float *light_response;        // scratch [n_vertices][nc][3]
  // Distinct (vertex, material) pairs of the triangle corners: corner c of
  // triangle t uses pair corner_key[3 t + c]; pair u is (key_vertex[u],
  // key_material[u]).
  const int32_t *corner_key, *key_vertex, *key_material;
  int64_t n_keys;
  float *key_color;             // scratch [n_keys][3]
  float cam_pos[3], cam_dir[3], cam_h[3], cam_v[3];
  float cx, cy, fx, fy;
  int width, height;
  const float *env_faces;       // view environment or null
  int env_n;
  float *color;                 // float4 [height][width]
  bool disney;                  // some material is DisneyBSDF
};
void launch_sh_render(const ShArgs &a, hipStream_t s);

// SH transfer matrices (gbuffer_kernels.hip): paths from the vertices
// [v_begin, v_begin + n_verts) traced once per (vertex, sample), their
// light samples and exit radiance under each of the nc basis maps, then
// the projection onto the filtered SH basis of the start directions.
constexpr int kTmMaxBounce = 15;
struct TmArgs {
  const BvhNode *nodes;
  const BvhTri *tris;
  const int32_t *indices;       // [n_tri][3]
  const float *vertices;        // [n_vertices][3]
  const float *normals;         // [n_vertices][3] or null
  const int32_t *tri_material;  // or null
  const MaterialRec *materials;
  const float *basis;           // [nc][6][n][n][3], nc = (degree + 1)^2
  const double *tabs;           // [nc][tab_doubles] sampling tables
  size_t tab_doubles;           // env_table_doubles(n)
  int n, degree, samples, max_bounce;
  uint32_t seed;
  long long v_begin, n_verts;
  void *work;                   // tm_work_bytes(n_verts, samples, degree, max_bounce, textured)
  float *transfer;              // [n_vertices][nc * nc][3]
  bool disney;                  // some material is DisneyBSDF
  const float *texcoords;       // as CastArgs (null: untextured)
  const TexRef *mat_tex;
};
// textured: the work also holds each path record's texture coordinates.
size_t tm_work_bytes(long long n_verts, int samples, int degree, int max_bounce, bool textured);
void launch_sh_transfer(const TmArgs &a, hipStream_t s);

// Persistent batch form: work item = the 64 pixels [base, base + 64) of job
// `job` (base < the job's pixel count).
struct BatchItem {
  int job, base;
};
// Whether a BF16/FP16 tile variant has a persistent batch form (output
// layer in VGPRs).
bool relight_batch_variant_ok(int variant);
// Waves take items [0, *n_items) in order from *counter, which must be 0 at
// launch (launch_reset_counter); the grid is n_cu x the variant's waves per
// SIMD workgroups.  Jobs have no transport tap and all use `softplus` as
// their output activation; has_aux: some job has aux inputs.
void launch_relight_batch(const RelightArgs *jobs, const BatchItem *items, const int *n_items,
                          unsigned int *counter, Elem elem, bool softplus, bool has_aux,
                          int variant, int n_cu, hipStream_t s);
void launch_reset_counter(unsigned int *counter, hipStream_t s);

// Host-side MFMA weight tiling of layers 2..4 and of the second layer's aux
// columns (see relight kernels).
void build_weight_tiles(const float *w_f32, Elem elem, std::vector<uint16_t> &out);
void build_aux_tiles(const float *waux, Elem elem, std::vector<uint16_t> &out);

// FP8 form (relight_fp8_kernels.hip).  Layers 2..4 quantized to OCP E4M3
// with one power-of-two scale per layer (amax mapped into (224, 448]);
// `exps` receives the three exponents.
void build_fp8_weight_tiles(const float *w_f32, std::vector<uint8_t> &out, int exps[3]);
// Calibration subsample of a frame of n_pixels: calibration pixel q is frame
// pixel q * stride, about kFp8CalibPixels of them.
constexpr int kFp8CalibPixels = 16384;
inline int fp8_calib_stride(int n_pixels) {
  return n_pixels <= kFp8CalibPixels ? 1 : (n_pixels + kFp8CalibPixels - 1) / kFp8CalibPixels;
}
inline int fp8_calib_count(int n_pixels) {
  const int s = fp8_calib_stride(n_pixels);
  return (n_pixels + s - 1) / s;
}
// Activation calibration: copies every `stride`-th pixel's BF16 features
// (zeros for background pixels) into `calib` [n_cal][32] (planar: the planar
// layout with n_cal pixels), and with `aux` its aux inputs into `calib_aux`
// [n_cal][16]; clears amax[4].
void launch_calib_gather(const uint16_t *features, const uint16_t *aux, const uint8_t *env_flag,
                         int n_pixels, int stride, uint16_t *calib, uint16_t *calib_aux,
                         int n_cal, float *amax, bool planar, hipStream_t s);
// The same calibration sample encoded from the G-buffer positions (`enc`:
// the fused-encode fields of a relight, its env_flag and n_pixels), planar
// BF16 features bit-identical to the standalone encode's; clears amax[4].
void launch_calib_encode(const RelightArgs &enc, int stride, uint16_t *calib, int n_cal,
                         float *amax, hipStream_t s);
// Maxima of the layer 1..3 activations of the BF16 network over the
// calibration pixels and all K wavelets (`a`: BF16 MFMA operands, features =
// calib, aux = calib_aux with waux_tiles, n_pixels = n_cal) into a.amax[0..2].
void launch_calibrate_amax(const RelightArgs &a, hipStream_t s);
int relight_fp8_variant_count();
// FP8 variant used when none was tuned or chosen (batches, servers, untuned
// shapes): the software-pipelined form.
constexpr int kDefaultFp8Variant = 1;
const char *relight_fp8_variant_name(int variant);
void launch_relight_fp8(const RelightArgs &a, int variant, hipStream_t s);
// Persistent FP8 batch form (items and counter as launch_relight_batch);
// has_aux: some job has aux inputs; variant: an FP8 variant (all have a
// persistent form).
void launch_relight_fp8_batch(const RelightArgs *jobs, const BatchItem *items, const int *n_items,
                              unsigned int *counter, bool softplus, bool has_aux, int variant,
                              int n_cu, hipStream_t s);
void build_fp8_aux_tiles(const float *waux, std::vector<uint16_t> &out);
// OCP E4M3 encode (round to nearest even, saturating at +-448) / decode.
uint8_t f32_to_e4m3(float x);
float e4m3_to_f32(uint8_t c);

}  // namespace nwprt
