/*
 * nwprt.h -- C ABI of the MI355X (gfx950) neural wavelet PRT relighting path.
 *
 * The reference renderer (AlstonXiao/neural_wavelet_prt_demo) has no FFI: its
 * relighting path is a set of C++ methods on `RenderEngine` plus on-disk
 * formats.  Every entry point below says which reference interface it
 * replaces (paths relative to the reference repository root).
 *
 * Conventions
 *  - Every function returns an nwprt_status; NWPRT_OK (0) is success.  On
 *    failure nwprt_last_error() returns a message for the calling thread.
 *    No C++ exception crosses this boundary.
 *  - "device" pointers are HIP device addresses on the context's GPU;
 *    "host" pointers are ordinary CPU memory.  The caller owns every buffer
 *    it passes in; the library owns scenes, environment maps and its own
 *    work buffers.
 *  - `stream` is a hipStream_t passed as void*; NULL selects the context's
 *    own stream.  Calls are asynchronous on that stream unless stated.
 *  - One context may be used from one host thread at a time.
 *  - Pixel buffers follow the reference G-buffer layout
 *    (Renderer/frames.h:143-183): float4 per pixel for positions, normals,
 *    reflection directions, kd, ks, direct light and outputs; one float per
 *    pixel for roughness; one byte per pixel for the environment-map flag.
 */
#ifndef NWPRT_H
#define NWPRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NWPRT_ABI_VERSION 1

/* Wavelet count and cubemap limits of this build. */
#define NWPRT_MAX_K 1024        /* selected wavelets per environment map */
#define NWPRT_MAX_FACE_SIZE 128 /* cubemap face edge N (power of two)    */
#define NWPRT_MAX_LEVELS 32     /* hash-grid levels                      */

typedef enum {
  NWPRT_OK = 0,
  NWPRT_ERR_INVALID_ARGUMENT = 1,
  NWPRT_ERR_IO = 2,           /* file missing / unreadable                   */
  NWPRT_ERR_FORMAT = 3,       /* malformed .npy / .hdr / json / shape        */
  NWPRT_ERR_DEVICE = 4,       /* HIP runtime error                           */
  NWPRT_ERR_OUT_OF_MEMORY = 5,
  NWPRT_ERR_UNSUPPORTED = 6,  /* valid input this build does not implement   */
  NWPRT_ERR_STATE = 7,        /* call order violated (e.g. p2 before p1)     */
  NWPRT_ERR_NOT_FOUND = 8     /* unknown scene / environment-map handle      */
} nwprt_status;

/* Arithmetic of the per-(pixel, wavelet) MLP. */
typedef enum {
  NWPRT_PREC_FP32 = 0, /* plain form: FP32 vector ALU, one pixel per lane    */
  NWPRT_PREC_FP16 = 1, /* fused MFMA, FP16 operands, FP32 accumulation       */
  NWPRT_PREC_BF16 = 2, /* fused MFMA, BF16 operands, FP32 accumulation       */
  NWPRT_PREC_FP8 = 3   /* fused MFMA: hidden layers FP8-E4M3 (OCP) with one
                          power-of-two scale per layer, weights from their
                          amax, activations from a BF16 calibration launch
                          over ~8K pixels of the frame; first layer BF16;
                          FP32 accumulation.  Scenes with aux inputs:
                          NWPRT_ERR_UNSUPPORTED. */
} nwprt_precision;

typedef enum { NWPRT_DTYPE_F32 = 0, NWPRT_DTYPE_F16 = 1 } nwprt_dtype;

/* Which float4 buffer the RGBA8 tone map reads
 * (Renderer/toneMap.cu:80-85: wavelet full -> with-direct, indirect -> indirect). */
typedef enum {
  NWPRT_TONEMAP_FULL = 0,
  NWPRT_TONEMAP_INDIRECT = 1,
  NWPRT_TONEMAP_NONE = 2
} nwprt_tonemap_source;

typedef enum { NWPRT_HOST = 0, NWPRT_DEVICE = 1 } nwprt_location;

typedef struct nwprt_context *nwprt_ctx;

/* ---------------------------------------------------------------- scene -- */

/* Multiresolution hash grid of the first-hit position (tiny-cuda-nn
 * "HashGrid" as configured by config_comp.json "vtx_enc",
 * Renderer/tinycudann/prt_network.h:45). */
typedef struct {
  int32_t n_levels;             /* 16  */
  int32_t n_features_per_level; /* 2   */
  int32_t log2_hashmap_size;    /* 19  */
  int32_t base_resolution;      /* 16  */
  float per_level_scale;        /* 1.5 */
} nwprt_hashgrid_desc;

/* Host-computed per-level table.  Computed once on the host in FP32 and
 * passed to kernels; never re-derived on the device. */
typedef struct {
  int32_t n_levels;
  float scale[NWPRT_MAX_LEVELS];        /* exp2(l*log2(b))*N_min - 1 (FP32) */
  uint32_t resolution[NWPRT_MAX_LEVELS];/* ceil(scale) + 1                  */
  uint32_t dense[NWPRT_MAX_LEVELS];     /* 1 if resolution^3 <= 2^log2T     */
  uint32_t size[NWPRT_MAX_LEVELS];      /* entries in the level             */
  uint32_t offset[NWPRT_MAX_LEVELS + 1];/* prefix sum of size; [n] = total  */
} nwprt_hashgrid_levels;

/* Network topology.  The per-pair network is
 *   z = proj(hash(pos)) (.) wavelet_feature[:, c_k]          (feature_dim)
 *   -> Linear(feature_dim, hidden) ReLU -> (n_hidden_layers-1) x
 *      [Linear(hidden, hidden) ReLU] -> Linear(hidden, 3) -> output act.
 * Bias-free, as tiny-cuda-nn networks are.  This build's fused kernels
 * implement feature_dim 32, hidden 64, 3 hidden layers. */
typedef struct {
  nwprt_hashgrid_desc grid;
  nwprt_dtype hash_dtype;    /* storage type of hash-grid entries          */
  int32_t face_size;         /* N: wavelet feature table is [F][6*N*N]      */
  int32_t feature_dim;       /* = n_levels * n_features_per_level (32)      */
  int32_t hidden_width;      /* 64                                          */
  int32_t n_hidden_layers;   /* 3                                           */
  int32_t has_projection;    /* 1: feature_dim x feature_dim projection     */
  int32_t output_softplus;   /* 1: softplus on the 3 outputs, 0: identity   */
  int32_t aux_dims;          /* 0, or 13: G-buffer encodings (normal,
                                reflect, roughness, kd, ks) entering the
                                second layer (tcnn.cu cmp_enc Identity)     */
  float aux_pad_value;       /* value of aux inputs 13..15 (padding)        */
} nwprt_scene_desc;

/* Host-side scene parameters (all row-major). */
typedef struct {
  const void *hash_table;        /* levels.offset[n] entries x F, hash_dtype */
  const float *wavelet_features; /* [feature_dim][6*N*N]                     */
  const float *projection;       /* [feature_dim][feature_dim] (out x in)    */
  const float *layers;           /* W1 [H][F], W2..Wn [H][H], Wout [3][H],
                                    then, if aux_dims, Waux [H][16]          */
} nwprt_scene_params;

/* ------------------------------------------------------------- G-buffer -- */

/* One frame's G-buffer (Renderer/frames.h:96-183, written by the OptiX pass
 * at Renderer/device_functions/camera.cu:247-270).  position: bbox-normalised
 * first hit in [0,1]^3; normal/reflect encoded as v/2+0.5; kd, ks albedos;
 * direct: first-hit direct light; env_flag: 1 where the primary ray hit
 * geometry.  All float4 except roughness (float) and env_flag (uint8). */
typedef struct {
  float *position;
  float *normal;
  float *reflect;
  float *roughness;
  float *kd;
  float *ks;
  float *direct;
  uint8_t *env_flag;
} nwprt_gbuffer;

/* Outputs (Renderer/frames.h:135-141): float4 (rgb, 1) per pixel and packed
 * RGBA8.  Any pointer may be NULL to skip that output. */
typedef struct {
  float *indirect;             /* fbPredictedIndirect              */
  float *indirect_with_direct; /* fbPredictedIndirectWithDirect    */
  uint32_t *rgba8;             /* fbFinalColor                     */
} nwprt_frame_out;

/* ---------------------------------------------------------------- core -- */

int32_t nwprt_abi_version(void);
const char *nwprt_last_error(void);
const char *nwprt_status_string(nwprt_status s);

/* Replaces the process-global TCNN state of Renderer/tinycudann/tcnn.cu:40-47
 * (network singleton + 4 streams) with an explicit context. */
nwprt_status nwprt_create(int device, nwprt_ctx *out);
nwprt_status nwprt_destroy(nwprt_ctx ctx);
/* RenderEngine::TCNNSync, Renderer/tinycudann/tcnn.cu:154-158. */
nwprt_status nwprt_sync(nwprt_ctx ctx);
/* The context's own stream (hipStream_t as void*). */
nwprt_status nwprt_get_stream(nwprt_ctx ctx, void **stream);

/* ------------------------------------------------------- host utilities -- */
/* These never touch the GPU. */

/* Hash-grid level table (tiny-cuda-nn grid conventions behind
 * Renderer/tinycudann/prt_network.h:45). */
nwprt_status nwprt_hashgrid_level_table(const nwprt_hashgrid_desc *desc,
                                        nwprt_hashgrid_levels *out);
/* Element counts of a scene's parameter arrays. */
nwprt_status nwprt_scene_param_counts(const nwprt_scene_desc *desc,
                                      size_t *hash_entries,
                                      size_t *wavelet_feature_count,
                                      size_t *projection_count,
                                      size_t *layer_weight_count);

/* Seeded synthetic inputs (counter-based SplitMix64 streams, Box-Muller in
 * double).  They stand in for the OptiX G-buffer pass and the offline
 * training / env-map data.  Outputs are host arrays sized by the caller. */
nwprt_status nwprt_synth_scene(const nwprt_scene_desc *desc, uint64_t seed,
                               float hash_std, float feature_std,
                               void *hash_table, float *wavelet_features,
                               float *projection, float *layers);
nwprt_status nwprt_synth_gbuffer(uint64_t seed, int32_t width, int32_t height,
                                 float background_fraction,
                                 const nwprt_gbuffer *host_out);
/* faces: [6][N][N][3] FP32 RGB, log-normal(0, 1.5) plus one 3x3 sun at 1e3. */
nwprt_status nwprt_synth_cubemap(uint64_t seed, int32_t face_size,
                                 float *faces);

/* File formats.  Arrays returned through `float **` / `int32_t **` are
 * allocated by the library and released with nwprt_free_host. */
void nwprt_free_host(void *p);
/* npy v1/v2/v3, dtypes <f4 <f8 <f2 <i4 <i8 <u1, C or Fortran order, converted
 * (as Renderer/npy.h:LoadArrayFromNumpy is used) to float32 or int32. */
nwprt_status nwprt_npy_read_f32(const char *path, float **data, int64_t *shape,
                                int32_t *ndim, int32_t *fortran_order);
nwprt_status nwprt_npy_read_i32(const char *path, int32_t **data,
                                int64_t *shape, int32_t *ndim,
                                int32_t *fortran_order);
nwprt_status nwprt_npy_write_f32(const char *path, const float *data,
                                 const int64_t *shape, int32_t ndim);
nwprt_status nwprt_npy_write_i32(const char *path, const int32_t *data,
                                 const int64_t *shape, int32_t ndim);
/* Radiance RGBE reader (Renderer/HDRloader.cpp:30-118): rows returned in
 * file order (top first), RGB float32. */
nwprt_status nwprt_hdr_read(const char *path, float **rgb, int32_t *width,
                            int32_t *height);
/* Radiance RGBE writer (flat, run-length free scanlines), rows top first. */
nwprt_status nwprt_hdr_write(const char *path, const float *rgb,
                             int32_t width, int32_t height,
                             int32_t channels);
/* 4N x 3N cubemap cross (rows top first, as read) -> faces [6][N][N][3] with
 * face order +X,-X,+Y,-Y,-Z,+Z and rows bottom-up
 * (Renderer/EnvironmentMap.cpp:112-126, Renderer/HDRloader.h:21-25,
 * Renderer/textures.cpp:72-83). */
nwprt_status nwprt_cubemap_cross_to_faces(const float *cross, int32_t width,
                                          int32_t height, float *faces);

/* ---------------------------------------------------------------- scene -- */

/* Upload a scene (hash grid, wavelet feature table, projection, MLP).
 * Replaces RenderEngine::initTCNN's upload half (tcnn.cu:67-74). */
nwprt_status nwprt_scene_create(nwprt_ctx ctx, const nwprt_scene_desc *desc,
                                const nwprt_scene_params *params,
                                int32_t *scene_id);
/* RenderEngine::initTCNN(modelPath), Renderer/tinycudann/tcnn.cu:49-81:
 * reads config_comp.json, vtx_enc.npy, out_prd.npy, cmp_net.npy, wve_enc.npy. */
nwprt_status nwprt_scene_load_folder(nwprt_ctx ctx, const char *model_dir,
                                     int32_t *scene_id);
nwprt_status nwprt_scene_destroy(nwprt_ctx ctx, int32_t scene_id);
nwprt_status nwprt_scene_get_desc(nwprt_ctx ctx, int32_t scene_id,
                                  nwprt_scene_desc *desc);

/* ---------------------------------------------------- environment maps -- */

/* Cubemap faces [6][N][N][3] FP32 -> GPU non-standard Haar transform (FP64
 * scaling-coefficient reduction) -> area-weighted top-K (ties: lower index).
 * Replaces the offline producer of waveletcoeffs/ and waveletstrength/
 * consumed at Renderer/EnvironmentMap.cpp:56-67. */
nwprt_status nwprt_envmap_from_faces(nwprt_ctx ctx, const float *faces,
                                     nwprt_location faces_location,
                                     int32_t face_size, int32_t k,
                                     int32_t *env_id);
/* Recompute an existing map (built from faces, same N and K) from new faces:
 * Haar + top-K into its buffers on `stream`, no allocation.  For requests
 * that carry a rotated or fresh cubemap. */
nwprt_status nwprt_envmap_update_faces(nwprt_ctx ctx, int32_t env_id,
                                       const float *faces,
                                       nwprt_location faces_location,
                                       void *stream);
/* No counterpart in the reference (a served request's rotated map).  As
 * nwprt_envmap_update_faces, from `faces` (device, [6][N][N][3], not the
 * map's own buffers) turned by `yaw` radians about +y: each texel takes the
 * source texel nearest to its rotated centre direction, with the face and
 * (u, v) conventions of EnvUVToDirec / EnvDirecToUV
 * (Renderer/device_functions/cubeMap.cu:23-115).  The first call on a map
 * allocates its rotated-faces buffer; later calls allocate nothing. */
nwprt_status nwprt_envmap_update_rotated(nwprt_ctx ctx, int32_t env_id,
                                         const float *faces, double yaw,
                                         void *stream);
/* Precomputed selection: indices int32[K] into [0, 6N^2) and strengths
 * float32[K][3] (Renderer/EnvironmentMap.cpp:56-67 +
 * RenderEngine::buildLights wavelet upload, Renderer/RenderEngine.cpp:167-175). */
nwprt_status nwprt_envmap_from_selection(nwprt_ctx ctx,
                                         const int32_t *indices,
                                         const float *strengths, int32_t k,
                                         int32_t face_size, int32_t *env_id);
/* Environment_Map(folder, name) (Renderer/EnvironmentMap.cpp:20-68):
 * cube/<name>.hdr is loaded; waveletcoeffs/<name>.npy + waveletstrength/
 * <name>.npy are used when both exist, otherwise the selection is computed
 * on the GPU with `k` wavelets. */
nwprt_status nwprt_envmap_load_folder(nwprt_ctx ctx, const char *folder,
                                      const char *name, int32_t k,
                                      int32_t *env_id);
/* Non-standard 2-D Haar transform of n_maps cubemaps: faces
 * [n_maps][6][N][N][3] FP32 in, coefficients in the same layout out (device
 * pointers, not overlapping), Mallat layout per face and channel, FP64
 * scaling coefficients, each output rounded to FP32 once -- the transform
 * whose flat indices the reference's waveletcoeffs/ files hold.  The
 * reference runs it offline and only loads its results
 * (Renderer/EnvironmentMap.cpp:36-40, 56-67); this is that producer, and the
 * first step of nwprt_envmap_from_faces.  Asynchronous on `stream`; it uses
 * a work buffer of the context, so calls on one context must not overlap on
 * different streams. */
nwprt_status nwprt_envmap_haar(nwprt_ctx ctx, const float *faces, int32_t face_size,
                               int32_t n_maps, float *coefficients, void *stream);

nwprt_status nwprt_envmap_destroy(nwprt_ctx ctx, int32_t env_id);
nwprt_status nwprt_envmap_info(nwprt_ctx ctx, int32_t env_id,
                               int32_t *face_size, int32_t *k);
/* Debug taps (synchronous): the selection (host int32[K], float[K][3]) and,
 * for maps built from faces, the FP32 Haar coefficients (host [6N^2][3]). */
nwprt_status nwprt_envmap_get_selection(nwprt_ctx ctx, int32_t env_id,
                                        int32_t *indices, float *strengths);
nwprt_status nwprt_envmap_get_coefficients(nwprt_ctx ctx, int32_t env_id,
                                           float *coefficients);

/* -------------------------------------------------------------- relight -- */

/* Whole relight of one frame: per-pixel encoding, per-pair MLP, clamped
 * compensated K-sum, outputs and tone map.  Replaces TCNNPass_p1 +
 * TCNNPass_p2 (+ computeFinalPixelColors when out->rgba8 is set),
 * Renderer/tinycudann/tcnn.cu:83-152, Renderer/toneMap.cu:73-99.
 * G-buffer and outputs are device pointers. */
nwprt_status nwprt_relight(nwprt_ctx ctx, int32_t scene_id, int32_t env_id,
                           const nwprt_gbuffer *gbuffer, int32_t width,
                           int32_t height, const nwprt_frame_out *out,
                           nwprt_precision precision,
                           nwprt_tonemap_source tonemap, void *stream);

/* Allocation-free relight.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * outputs as nwprt_relight -- equal bit for bit -- for BF16, FP16 and FP8 on
 * scenes with a 16 x 2 hash grid and no aux inputs (NWPRT_ERR_UNSUPPORTED
 * otherwise).  Its work buffers live in a 10 MiB array that each device
 * holds until the process ends (enough for any K <= 1024), so it calls no
 * hipMalloc or hipFree.  Calls that overlap on one GPU share that array and
 * race: serialise them. */
nwprt_status nwprt_relight_noalloc(nwprt_ctx ctx, int32_t scene_id, int32_t env_id,
                                   const nwprt_gbuffer *gbuffer, int32_t width,
                                   int32_t height, const nwprt_frame_out *out,
                                   nwprt_precision precision,
                                   nwprt_tonemap_source tonemap, void *stream);

/* Split form for a two-frame pipeline (RenderEngine::TCNNPass_p1 /
 * TCNNPass_p2, tcnn.cu:83-152; OutputVisualizer.cpp:61-126).  p1 encodes the
 * environment-independent per-pixel features of `frame_slot` (0 or 1); p2
 * runs the per-pair network for an environment map on that slot.  p2 on a
 * slot not encoded at this size/scene/precision returns NWPRT_ERR_STATE
 * (the reference asserts, tcnn.cu:126). */
nwprt_status nwprt_pass_p1(nwprt_ctx ctx, int32_t scene_id, int32_t frame_slot,
                           const nwprt_gbuffer *gbuffer, int32_t width,
                           int32_t height, nwprt_precision precision,
                           void *stream);
nwprt_status nwprt_pass_p2(nwprt_ctx ctx, int32_t scene_id, int32_t env_id,
                           int32_t frame_slot, const nwprt_gbuffer *gbuffer,
                           int32_t width, int32_t height,
                           const nwprt_frame_out *out,
                           nwprt_precision precision,
                           nwprt_tonemap_source tonemap, void *stream);

/* ------------------------------------------------------- tile autotuning -- */
/* No counterpart in the reference (its MLP tiling is fixed inside
 * tiny-cuda-nn).  The fused MFMA kernel comes in several tile-shape
 * variants (pixel groups per pass, output-layer weights in VGPRs or LDS,
 * waves per SIMD).  nwprt_autotune_relight times every variant on one frame
 * and records the fastest for that (precision, K, pixel-count) bucket;
 * later relights falling in the bucket use it, others use variant 0.
 * The count and names below are those of the BF16 / FP16 kernel; the FP8
 * kernel has its own (nwprt_relight_fp8_variant_*), and autotuning or
 * querying with NWPRT_PREC_FP8 works on those. */
int32_t nwprt_relight_variant_count(void);
const char *nwprt_relight_variant_name(int32_t variant);
int32_t nwprt_relight_fp8_variant_count(void);
const char *nwprt_relight_fp8_variant_name(int32_t variant);
/* Force one BF16 / FP16 variant for every bucket (-1 returns to the tuned
 * choice); FP8 relights use it too when it also names an FP8 variant. */
nwprt_status nwprt_set_relight_variant(nwprt_ctx ctx, int32_t variant);
nwprt_status nwprt_autotune_relight(nwprt_ctx ctx, int32_t scene_id,
                                    int32_t env_id,
                                    const nwprt_gbuffer *gbuffer,
                                    int32_t width, int32_t height,
                                    nwprt_precision precision,
                                    int32_t *best_variant, float *best_ms);
/* The variant a relight of this shape would use now. */
nwprt_status nwprt_selected_relight_variant(nwprt_ctx ctx,
                                            nwprt_precision precision,
                                            int32_t k, int32_t n_pixels,
                                            int32_t *variant);

/* RenderEngine::computeFinalPixelColors kernel (Renderer/toneMap.cu:47-71):
 * float4 -> RGBA8, device pointers. */
nwprt_status nwprt_tonemap(nwprt_ctx ctx, const float *src_float4,
                           uint32_t *dst_rgba8, int32_t width, int32_t height,
                           void *stream);

/* -------------------------------------------------------------- denoiser -- */
/* OptixBackend::denoise(color, normal, albedo, output, blendFactor, size,
 * stream) (Renderer/OptixBackend.cpp:841-906), as called by
 * RenderEngine::denoiseOrMoveData with the denoiser on
 * (Renderer/RenderEngine.cpp:390-397) and by TCNNPass_p2 on its two outputs
 * (Renderer/tinycudann/tcnn.cu:144-148).  All images are device float4
 * [height][width] (the reference's OptixImage2D FLOAT4 layout); the normal
 * and albedo buffers are the guides (guideNormal / guideAlbedo,
 * OptixBackend.cpp:783-785; world or camera normals give the same result,
 * only differences between normals enter).  The reference's learned model
 * is replaced by an edge-avoiding a-trous wavelet filter:
 * - kappa = 2^-mean(log2 Y) over pixels with luminance Y > 1e-8 (the role
 *   of optixDenoiserComputeIntensity, OptixBackend.cpp:871-877);
 * - `iterations` passes, step 2^i, 5x5 B3-spline taps, weights
 *   exp(-(kappa^2 |dc|^2 4^i / sigma_color^2 + |dn|^2 / sigma_normal^2 +
 *   |da|^2 / sigma_albedo^2)), taps outside the image skipped;
 * - RGB = blend * color + (1 - blend) * filtered; alpha copied from color
 *   (OPTIX_DENOISER_ALPHA_MODE_COPY, OptixBackend.cpp:815).
 * `output` must not overlap the three inputs.  Work buffers (2 images) are
 * held by the context, like the reference's per-size denoiser state
 * (OptixBackend::setupDenoiser, OptixBackend.cpp:772-817).  params NULL
 * selects the defaults below. */
typedef struct {
  int32_t iterations;  /* 1..8 a-trous passes (default 5)                  */
  float sigma_color;   /* colour scale, kappa-normalised (default 1.0)     */
  float sigma_normal;  /* normal scale (default 0.3)                       */
  float sigma_albedo;  /* albedo scale (default 0.1)                       */
} nwprt_denoise_params;

void nwprt_denoise_default_params(nwprt_denoise_params *params);
nwprt_status nwprt_denoise(nwprt_ctx ctx, const float *color, const float *normal,
                           const float *albedo, float *output, int32_t width,
                           int32_t height, float blend_factor,
                           const nwprt_denoise_params *params, void *stream);

/* ---------------------------------------------------- G-buffer producer -- */
/* The first-hit part of the OptiX pass that fills the G-buffer before the
 * relighting path (__raygen__fovCamera, Renderer/device_functions/
 * camera.cu:82-290; __closesthit__radiance / __miss__radiance,
 * Renderer/device_functions/devicePrograms.cu:47-311, 517-564), as a HIP
 * ray caster over a BVH resident in device memory.  It writes the same
 * encodings the relighting path reads. */

/* Camera::from / at / up as set by RenderEngine::setCamera
 * (Renderer/RenderEngine.cpp:246-267): 0.8576 rad field of view, principal
 * point at the image centre. */
typedef struct {
  float from[3];
  float at[3];
  float up[3];
} nwprt_camera;

/* A material as Renderer/Material.h:21-117 holds it (constant values; no
 * texture maps): kd (DisneyBSDF's base colour), ks, roughness, the
 * Material::MatType (Renderer/Material.h:7: 0 roughPlastic, 1 DisneyBSDF,
 * 2 cookTorrance, 3 mirror, which reflects the ray instead of ending it,
 * 4 simple_ggx), the DisneyBSDF parameters and the index of refraction eta
 * that roughPlastic and DisneyBSDF use.  nwprt_material_defaults fills the
 * reference's defaults (Material::init and the per-type constructors,
 * Material.h:79-155). */
typedef struct {
  float kd[3];
  float ks[3];
  float roughness;
  int32_t type;
  float specular_transmission;
  float metallic;
  float subsurface;
  float specular;
  float specular_tint;
  float anisotropic;
  float sheen;
  float sheen_tint;
  float clearcoat;
  float clearcoat_gloss;
  float eta;
} nwprt_material;

/* Host utility (no GPU): the reference's default material of a type. */
nwprt_status nwprt_material_defaults(int32_t type, nwprt_material *out);

/* Host utility (no GPU): an 8-bit material texture as
 * EightBitTexture::loadTexture reads it (Renderer/textures.cpp:102-127,
 * through stb_image): RGBA8, row 0 at the bottom of the picture.  PNG (all
 * colour types and bit depths, palette / tRNS, Adam7), TGA (true-colour or
 * grey, raw or RLE), BMP (24/32-bit uncompressed), binary PGM / PPM; grey
 * expands to (g, g, g), missing alpha is 255, 16-bit samples keep their
 * high byte.  Other formats (JPEG among them) -> NWPRT_ERR_UNSUPPORTED.
 * *rgba [height][width][4] is released with nwprt_free_host. */
nwprt_status nwprt_image_read_rgba8(const char *path, uint8_t **rgba, int32_t *width,
                                    int32_t *height);

#define NWPRT_MATERIAL_MIRROR 3

/* Material texture fields (the SBT fields devicePrograms.cu:121-147 reads
 * from textures when the mesh has texture coordinates): kd and ks take the
 * texel's RGB, the others its first channel. */
#define NWPRT_TEX_KD 0
#define NWPRT_TEX_KS 1
#define NWPRT_TEX_ROUGHNESS 2
#define NWPRT_TEX_SPECULAR_TRANSMISSION 3
#define NWPRT_TEX_METALLIC 4
#define NWPRT_TEX_SUBSURFACE 5
#define NWPRT_TEX_SPECULAR 6
#define NWPRT_TEX_SPECULAR_TINT 7
#define NWPRT_TEX_ANISOTROPIC 8
#define NWPRT_TEX_SHEEN 9
#define NWPRT_TEX_SHEEN_TINT 10
#define NWPRT_TEX_CLEARCOAT 11
#define NWPRT_TEX_CLEARCOAT_GLOSS 12
#define NWPRT_TEX_ETA 13
#define NWPRT_TEX_FIELDS 14

typedef struct {
  float lower[3], upper[3]; /* bounds of the vertices (Scene::bounds)       */
  int64_t triangles;
  int64_t bvh_nodes;
  int32_t bvh_depth;        /* internal levels, <= 24                        */
  int64_t vertices;
} nwprt_mesh_info;

/* Uploads a triangle mesh and builds its BVH on the host (binned SAH, at
 * most 8 triangles per leaf, at most 24 internal levels, so at most 2^27
 * triangles); replaces the
 * OptiX GAS build (OptixBackend.cpp buildAccel).  vertices [n_vertices][3];
 * normals [n_vertices][3] or NULL (geometric normals); indices
 * [n_triangles][3]; material_ids [n_triangles] or NULL (all material 0);
 * materials [n_materials].  Host arrays; vertices are in world space. */
nwprt_status nwprt_mesh_create(nwprt_ctx ctx, const float *vertices, int64_t n_vertices,
                               const float *normals, const int32_t *indices,
                               int64_t n_triangles, const int32_t *material_ids,
                               const nwprt_material *materials, int32_t n_materials,
                               int32_t *mesh_id);
nwprt_status nwprt_mesh_get_info(nwprt_ctx ctx, int32_t mesh_id, nwprt_mesh_info *out);
/* Host utility (no GPU): Wavefront OBJ (+ MTL) reader with what loadOBJ
 * (Renderer/Model.cpp:73-146) keeps: polygons fanned into triangles,
 * materials as cookTorrance (type 2: kd = Kd, ks = Ks, roughness =
 * (1 / Ns)^2, 0 for Ns = 0; faces without a material: kd = ks = 0.5,
 * roughness 1).  Vertices in the reference's scene order: one mesh per
 * (o / g shape, material), materials ascending, each mesh's distinct
 * (position, normal, texcoord) corners by first use (addVertex,
 * Model.cpp:34-70), the meshes one after the other.  *normals is NULL when
 * the file has none.  Arrays are released with nwprt_free_host. */
nwprt_status nwprt_obj_read(const char *path, float **vertices, int64_t *n_vertices,
                            float **normals, int32_t **indices, int64_t *n_triangles,
                            int32_t **material_ids, nwprt_material **materials,
                            int32_t *n_materials);
/* Host utility (no GPU): nwprt_obj_read plus the texture coordinates
 * ([n_vertices][2], NULL when the file has none) and, per material, the
 * files of its kd, ks and roughness maps (map_Kd, map_Ks, map_Ns;
 * Model.cpp:133-135) resolved against the OBJ's folder: *map_paths holds
 * n_materials x 3 NUL-terminated strings, "" for no map (*map_bytes in
 * all).  The reference loads a map only when its file exists
 * (textureManager::addTexture, textures.cpp:129-152). */
nwprt_status nwprt_obj_read_textured(const char *path, float **vertices, int64_t *n_vertices,
                                     float **normals, float **texcoords, int32_t **indices,
                                     int64_t *n_triangles, int32_t **material_ids,
                                     nwprt_material **materials, int32_t *n_materials,
                                     char **map_paths, int64_t *map_bytes);
/* Host utility (no GPU): Stanford PLY reader (ASCII or binary little-endian)
 * with what loadPly (Renderer/Model.cpp:152-220) keeps: vertex x y z,
 * optional nx ny nz, face vertex_indices (polygons fanned).  One material:
 * cookTorranceMaterial's defaults (Model.cpp:231-235).  Same outputs as
 * nwprt_obj_read. */
nwprt_status nwprt_ply_read(const char *path, float **vertices, int64_t *n_vertices,
                            float **normals, int32_t **indices, int64_t *n_triangles,
                            int32_t **material_ids, nwprt_material **materials,
                            int32_t *n_materials);
nwprt_status nwprt_mesh_destroy(nwprt_ctx ctx, int32_t mesh_id);

/* A material texture in device memory (EightBitTexture::toCUDA,
 * Renderer/textures.cpp:8-47 and textures.h:74-90): RGBA8 texels [height]
 * [width][4] with row 0 at v = 0 (as nwprt_image_read_rgba8 returns them),
 * sampled like the reference's CUDA texture object: normalized coordinates
 * wrapped in both directions, bilinear filtering between texel centres
 * ((i + 0.5) / width), bytes read as b / 255 with the sRGB decode on the
 * colour channels (tex_desc.sRGB is set), no mipmaps.  Host pixels,
 * copied. */
nwprt_status nwprt_texture_create(nwprt_ctx ctx, const uint8_t *rgba, int32_t width,
                                  int32_t height, int32_t *texture_id);
/* NWPRT_ERR_STATE while a mesh uses the texture. */
nwprt_status nwprt_texture_destroy(nwprt_ctx ctx, int32_t texture_id);
/* Texture coordinates and material textures of a mesh (the SBT's texcoord
 * array and per-field hasTexture / texture, OptixBackend.cpp:7-8):
 * texcoords [n_vertices][2] (host; NULL: none, and then, as without
 * hasTexcoord, no texture is read); material_textures [n_materials]
 * [NWPRT_TEX_FIELDS] texture ids, -1 for the material's constant (NULL:
 * none).  At a hit the coordinates are interpolated with the barycentrics,
 * (1 - u - v) t0 + u t1 + v t2, and each textured field replaces the
 * material's value in the first-hit encodings, the direct term and the
 * transfer computation.  Replaces earlier textures of the mesh; the mesh
 * holds its textures until it is destroyed or its textures are replaced.
 * The SH render does not take a textured mesh (NWPRT_ERR_UNSUPPORTED): the
 * reference reads its textures per pixel, which its per-vertex colours
 * cannot represent. */
nwprt_status nwprt_mesh_set_textures(nwprt_ctx ctx, int32_t mesh_id, const float *texcoords,
                                     const int32_t *material_textures);

typedef struct {
  int32_t max_bounce;          /* samplingParameter.maximumBounce: 1 in
                                  wavelet mode (RenderEngine.cpp:343-346);
                                  mirrors extend the path up to this depth */
  const float *bounds;         /* host float[6], lower then upper, used to
                                  normalise positions, or NULL for the
                                  mesh's own bounds                         */
  const float *view_env_faces; /* device [6][n][n][3] FP32 cube faces seen
                                  by primary rays that miss (kd = albedo =
                                  radiance), or NULL (zero)                 */
  int32_t view_env_n;
  const float *light_env_faces;/* device [6][n][n][3] lighting environment
                                  for the direct term, or NULL: `direct` is
                                  then not written                          */
  int32_t light_env_n;         /* 1..128                                    */
  int32_t direct_samples;      /* estimates averaged per pixel, each one
                                  environment sample + one BSDF sample      */
  uint32_t seed;               /* random stream of the direct term          */
} nwprt_cast_params;

/* One primary ray per pixel through the pixel's sample position at the
 * jitter's mean (screen = ((ix - cx) / fx, (iy - cy) / fy)).  Writes, per
 * pixel of a device G-buffer (any field may be NULL to skip it):
 * - hit: position (p + 0.001 - lower) / (upper - lower) with p the first hit,
 *   normal n / 2 + 0.5 (interpolated vertex normal, else geometric),
 *   reflect r / 2 + 0.5 (r = reflect(ray, n)), kd, ks, roughness of its
 *   material, env_flag 1;
 * - miss: position at (-1e5 - lower) / (upper - lower), normal and reflect
 *   -ray / 2 + 0.5, kd = view environment radiance (0 after a mirror),
 *   ks 0, roughness 1, env_flag 0;
 * - a mirror path cut by max_bounce: position -lower / (upper - lower),
 *   normal and reflect 0.5, kd, ks, roughness 0, env_flag 0.
 * The w components are 1.
 * direct (when light_env_faces is given and gbuffer->direct is not NULL):
 * the closest-hit program's direct term in wavelet mode
 * (devicePrograms.cu:165-280): per sample, one environment-importance
 * sample (face and texel tables of EnvironmentMap.cpp:153-190) and one BSDF
 * sample, each shadow-tested and weighted by the power heuristic, averaged
 * over direct_samples; misses store the view radiance, as kd.  Materials
 * may be roughPlastic (type 0, with the material's eta), DisneyBSDF (1,
 * disneyBSDF.cu: diffuse, subsurface blend, sheen, metal, clearcoat,
 * glass), cookTorrance (2), mirrors (3) or simple_ggx (4).  Random numbers come from a counter-based generator keyed by
 * (seed, pixel, sample), not the reference's per-pixel LCG: estimates agree
 * with the reference's in expectation, not sample by sample.  The sampling
 * tables live in a work buffer of the context: casts with a lighting map
 * issued on different streams must not overlap. */
nwprt_status nwprt_gbuffer_cast(nwprt_ctx ctx, int32_t mesh_id, const nwprt_camera *camera,
                                int32_t width, int32_t height,
                                const nwprt_cast_params *params,
                                const nwprt_gbuffer *gbuffer, void *stream);

/* ---------------------------------------------------------------- SH-PRT -- */
/* RenderEngine::shRender (Renderer/RenderEngine.cpp:459-486) with
 * shLoadTMatrix's transfer matrices (:448-457) and an environment's SH
 * coefficients (Environment_Map::load_sh, EnvironmentMap.cpp:71-78): the
 * SH-PRT baseline beside wavelet PRT.  For the first hit of each pixel's
 * primary ray (as nwprt_gbuffer_cast casts it), each of the hit triangle's
 * vertices gets |sum_k R_k * B_k| with R = T_v L (T_v the vertex's
 * [nc][nc][3] transfer matrix, row k, column j; L the [nc][3] light
 * vector; nc = (degree + 1)^2) and B the Monte Carlo projection of the
 * BSDF onto pbrt's real SH basis at the vertex (vertex normal, direction to
 * the camera, brdf_samples BSDF samples weighted f / pdf,
 * devicePrograms.cu:312-501); the pixel interpolates the three colours
 * barycentrically.  Misses show the view environment (or 0).  Vertex
 * colours are computed once per distinct (vertex, material) pair of the
 * triangle corners with a random stream keyed by (seed, vertex, material),
 * not once per pixel: the same expectation.
 * transfer: device [n_vertices][nc][nc][3]; light_sh: device [nc][3];
 * color: device float4 [height][width].  The mesh needs vertex normals;
 * materials must be roughPlastic, DisneyBSDF, cookTorrance or simple_ggx
 * (the reference's SH program calls the material's own callables; it has
 * none for mirrors).  The light
 * responses and vertex colours live in work buffers of the context: calls
 * issued on different streams must not overlap. */
nwprt_status nwprt_sh_render(nwprt_ctx ctx, int32_t mesh_id, const nwprt_camera *camera,
                             int32_t width, int32_t height, const float *transfer,
                             const float *light_sh, int32_t degree, int32_t brdf_samples,
                             uint32_t seed, const float *view_env_faces, int32_t view_env_n,
                             float *color, void *stream);

/* RenderEngine::shComputeTMatrix (Renderer/RenderEngine.cpp:410-446) with
 * its ray generation program __raygen__ComputeTMatrix
 * (Renderer/device_functions/cameraShTmatrix.cu:87-178): the transfer
 * matrices nwprt_sh_render reads, for the vertices [vertex_begin,
 * vertex_begin + vertex_count) (vertex_count -1: to the last vertex).
 * basis_faces: device [nc][6][face_n][face_n][3] FP32, the nc = (degree+1)^2
 * basis maps in order (SHRenderer.cpp:12-22 loads them from a folder of
 * .npy crosses in natural order); map i is at once the view, lighting and
 * sampling environment of its pass.  Per vertex, `samples` paths (4096 in
 * the reference) start at the vertex position in uniform sphere directions
 * and follow __closesthit__radiance / __miss__radiance
 * (devicePrograms.cu:47-311, 517-564): one environment-importance light
 * sample per non-mirror hit (power heuristic; max_bounce 1 adds the BSDF
 * light sample), BSDF continuations, mirrors reflect, Russian roulette after
 * depth 3, at most max_bounce rays (12 in the reference, 1..15 here); the
 * environment radiance that reached the vertex on the first ray, or on the
 * second after a mirror, is removed.  transfer: device
 * [n_vertices][nc * nc][3], row k * nc + i = SH coefficient k (of the
 * filtered basis 1 / (1 + 0.005 l^2 (l+1)^2) Y_k at the start direction)
 * of the indirect radiance under basis map i, times 4 pi / samples; rows of
 * other vertices are left alone.  Random numbers come from the
 * counter-based generator keyed by (seed, vertex), the same for every
 * basis map, as the reference's per-vertex stream is: estimates agree with
 * the reference's in expectation, not sample by sample.  Materials:
 * roughPlastic, DisneyBSDF, cookTorrance, mirror, simple_ggx.  The sampling
 * tables and path records live in work buffers of the context: calls issued
 * on different streams must not overlap. */
nwprt_status nwprt_sh_compute_transfer(nwprt_ctx ctx, int32_t mesh_id, const float *basis_faces,
                                       int32_t face_n, int32_t degree, int32_t samples,
                                       int32_t max_bounce, uint32_t seed, int64_t vertex_begin,
                                       int64_t vertex_count, float *transfer, void *stream);

/* ------------------------------------------------------------ debug taps -- */

/* Hash-grid tap: for n positions (device float4), write the 8 corner entry
 * indices per level (device uint32[n][L][8], index within the level) and
 * the interpolated features (device float[n][L*F]).  Either output may be
 * NULL. */
nwprt_status nwprt_debug_hash_encode(nwprt_ctx ctx, int32_t scene_id,
                                     const float *positions, int32_t n,
                                     uint32_t *indices, float *features,
                                     void *stream);
/* Transport tap: per-pair network output T (after output activation) for
 * pixels [pixel_begin, pixel_begin + n_pixels) of a frame: device
 * float[n_pixels][K][3].  Uses the same kernels as nwprt_relight. */
nwprt_status nwprt_debug_transport(nwprt_ctx ctx, int32_t scene_id,
                                   int32_t env_id, const nwprt_gbuffer *gbuffer,
                                   int32_t width, int32_t height,
                                   int32_t pixel_begin, int32_t n_pixels,
                                   float *transport, nwprt_precision precision,
                                   void *stream);

/* -------------------------------------------------------------- batching -- */
/* No counterpart in the reference, which relights one frame per call
 * (RenderEngine::TCNNPass_p2, Renderer/tinycudann/tcnn.cu:111-152).  A batch
 * is a fixed set of relight jobs -- each its own scene, environment map (and
 * so K), frame size, G-buffer and outputs, all device pointers -- run as one
 * captured HIP graph: per job the hash-grid encode and the selection tiles,
 * then ONE persistent fused kernel whose waves pull 64-pixel work items of
 * every job (largest K first) from a shared counter.  Each job's outputs
 * equal, bit for bit, those of nwprt_relight with the same precision and
 * tile variant.  The graph reads the G-buffers, environment-map selections
 * and outputs through the pointers given here, so new contents (e.g. an
 * nwprt_envmap_update_faces of a job's map, same K) are picked up by the
 * next launch.  BF16 / FP16 only; scenes without aux inputs; all jobs must
 * share the output activation (NWPRT_ERR_UNSUPPORTED otherwise). */
typedef struct {
  int32_t scene_id;
  int32_t env_id;
  nwprt_gbuffer gbuffer;
  int32_t width, height;
  nwprt_frame_out out;
} nwprt_batch_job;

nwprt_status nwprt_batch_create(nwprt_ctx ctx, const nwprt_batch_job *jobs,
                                int32_t n_jobs, nwprt_precision precision,
                                nwprt_tonemap_source tonemap,
                                int32_t *batch_id);
/* Launches the batch's graph on `stream` (NULL: the context stream) after
 * the jobs' environment maps are ready.  Asynchronous. */
nwprt_status nwprt_batch_launch(nwprt_ctx ctx, int32_t batch_id, void *stream);
/* Times every tile variant that has a persistent form on this batch and
 * keeps the fastest (re-captures the graph).  Synchronous. */
nwprt_status nwprt_batch_autotune(nwprt_ctx ctx, int32_t batch_id,
                                  int32_t *best_variant, float *best_ms);
/* Work items (64-pixel groups), (pixel, wavelet) pairs, tile variant. */
nwprt_status nwprt_batch_info(nwprt_ctx ctx, int32_t batch_id,
                              int64_t *n_items, int64_t *n_pairs,
                              int32_t *variant);
nwprt_status nwprt_batch_destroy(nwprt_ctx ctx, int32_t batch_id);

/* --------------------------------------------------------------- serving -- */
/* No counterpart in the reference (an interactive viewer relights one frame
 * at a time, Renderer/OutputVisualizer.cpp:61-126).  A server takes relight
 * requests as they arrive and keeps the GPU busy with all of them:
 *  - submit() queues a request (scene, resident G-buffer, outputs, and an
 *    environment map that is resident, a yaw rotation of a resident cubemap,
 *    or a fresh cubemap);
 *  - poll() retires finished work, admits queued requests in arrival order
 *    while the admitted ones' device state fits the memory budget and a slot
 *    is free, prepares each admitted request on a second stream (cubemap
 *    rotation or upload, Haar + top-K, hash-grid encode, selection tiles),
 *    and keeps up to two persistent relight launches ("ticks") in flight,
 *    each filled in arrival order with up to tick_pairs (pixel, wavelet)
 *    pairs of 64-pixel work items from every admitted request.
 * A request that arrives while others run joins the next tick; its
 * preparation overlaps their launches.  Each request's outputs equal, bit
 * for bit, nwprt_relight of the same inputs with the same precision and tile
 * variant.  Requests never share outputs while in flight (the caller's
 * responsibility); G-buffers and source cubemaps must stay valid until the
 * request completes, and the library keeps the request's scene and resident
 * map from being destroyed or updated until then.  A resident map's
 * strengths are copied at preparation. */
typedef struct {
  uint64_t memory_budget_bytes; /* device bytes of admitted requests' state; 0: no limit */
  int32_t max_active;           /* requests admitted at once (>= 1)         */
  int64_t tick_pairs;           /* pairs per persistent launch; 0: 2^30     */
  nwprt_precision precision;    /* BF16 or FP16                             */
  nwprt_tonemap_source tonemap;
  int32_t variant;              /* tile variant with a persistent form (FP8: an FP8
                                   variant), -1: default */
} nwprt_server_config;

typedef enum {
  NWPRT_ENV_RESIDENT = 0, /* env_id: an existing map (its K)                   */
  NWPRT_ENV_ROTATED = 1,  /* faces (device) turned by yaw: Haar + top-K of K   */
  NWPRT_ENV_FRESH = 2     /* faces (host or device): Haar + top-K of K         */
} nwprt_env_source;

typedef struct {
  int32_t scene_id;
  nwprt_gbuffer gbuffer;
  int32_t width, height;
  nwprt_frame_out out;
  nwprt_env_source env_source;
  int32_t env_id;
  const float *faces;
  nwprt_location faces_location;
  int32_t face_size;
  int32_t k;
  double yaw;
} nwprt_request;

typedef struct {
  int64_t submitted, completed, queued, active;
  int64_t ticks, items;
  uint64_t bytes_in_use, peak_bytes;
  int32_t peak_active;
  int64_t failed; /* requests reported done with a negative latency */
} nwprt_server_stats;

nwprt_status nwprt_server_create(nwprt_ctx ctx, const nwprt_server_config *config,
                                 int32_t *server_id);
nwprt_status nwprt_server_submit(nwprt_ctx ctx, int32_t server_id,
                                 const nwprt_request *request,
                                 int64_t *request_id);
/* Advances the server and reports up to max_done completed requests (ids and
 * submit-to-completion latency in ms, in completion order).  wait: 0 returns
 * at once; 1 blocks until at least one request completes (or none is
 * pending); 2 blocks until every submitted request has completed.  A request
 * that could not be run (its device state could not be allocated with no
 * other request admitted, or its preparation failed) is reported done with
 * latency -status (negative); it gives back everything it held, and the
 * queue moves on.  Scenes and resident maps that a queued or running
 * request reads cannot be destroyed or updated in place until it is
 * reported (NWPRT_ERR_STATE). */
nwprt_status nwprt_server_poll(nwprt_ctx ctx, int32_t server_id, int32_t wait,
                               int64_t *done_ids, double *latency_ms,
                               int32_t max_done, int32_t *n_done);
nwprt_status nwprt_server_get_stats(nwprt_ctx ctx, int32_t server_id,
                                    nwprt_server_stats *stats);
nwprt_status nwprt_server_destroy(nwprt_ctx ctx, int32_t server_id);

#ifdef __cplusplus
}
#endif

#endif /* NWPRT_H */
