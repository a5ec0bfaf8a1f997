/*
 * oracle.h -- CPU restatement of the neural wavelet PRT relighting path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.  The
 * product (libnwprt) never links, calls or falls back to it.
 *
 * Arithmetic is FP64 except where the reference path fixes an FP32 result
 * that later steps depend on bit for bit (hash-grid positions, scores of the
 * top-K selection, the tone map's byte quantisation); those follow the FP32
 * operation order stated at each function.  Citations are to the reference
 * repository (paths relative to its root).
 *
 * Parity status of each function (how it is pinned independently of itself):
 *   orc_hashgrid_levels  committed level-table fixture + FP64 evaluation
 *                        (tests/test_oracle_hashgrid.py)
 *   orc_hash_encode      brute-force trilinear interpolation of a dense grid,
 *                        one-hot weights on grid vertices, the published
 *                        spatial hash on hand-computed inputs
 *   orc_haar_face        closed forms (constant face, delta), Parseval,
 *                        inverse round trip, pywt-free textbook recursion
 *   orc_topk             full sort on (score desc, index asc), ties, K = all
 *   orc_rotate_cubemap   yaw 0 and 4 x 90 degrees = identity, 90 degrees =
 *                        the face permutation the axes give on face-constant
 *                        maps, 90 + 90 = 180, the +-Y faces turn in place
 *   orc_relight_pixels   brute-force numpy FP64 network on tiny cases,
 *                        closed forms (zero weights -> softplus(0) = ln 2,
 *                        background -> 0, negative light -> clamped)
 *   orc_tonemap          closed forms (0 -> 0, 1 -> 255, alpha 255, byte order);
 *                        every threshold checked against the FP64 curve in
 *                        Python (it reaches b / 255.9, the float below does
 *                        not); the reference's FP32 formula in numpy agrees
 *                        away from the thresholds
 *   orc_denoise          constant image and blend 1 unchanged; unbounded sigmas =
 *                        normalised B3-spline a-trous convolution (scipy); a
 *                        guide edge stops all flow across it; power-of-two
 *                        colour scaling commutes; transpose symmetry; the
 *                        learned model it replaces is closed, so parity with
 *                        the reference's denoised images is out of reach
 *   orc_gbuffer_cast     closed forms (plane at a known depth, centre-pixel ray
 *                        = view direction, mirror at 45 degrees), an
 *                        independent numpy edge-function intersection on
 *                        random triangle soups, cube-face lookups at texel
 *                        centres and axis directions, translation invariance
 *   orc_gbuffer_cast_direct  table invariants and inverse-CDF round trips;
 *                        the estimator's mean over many samples against
 *                        an independent quadrature of its expectation
 *                        (diffuse cookTorrance, no occluders); closed forms
 *                        of the BSDF (white furnace bound, reciprocity)
 *   orc_sh_eval          orthonormality by quadrature; closed forms of
 *                        bands 0-2 (Condon-Shortley phase)
 *   orc_sh_render        a single-entry transfer matrix (orientation) with
 *                        a diffuse BSDF against quadrature of the BSDF's
 *                        albedo
 * The tiny-cuda-nn conventions behind the hash grid and the MLP (level
 * rounding, layer layout, padding values) cannot be pinned against the
 * reference: that library is an empty submodule there ("parity unpinned"
 * for those conventions; see the design notes).
 */
#ifndef NWPRT_ORACLE_H
#define NWPRT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int n_levels, n_features;   /* L, F; L * F = feature_dim           */
  const float *scale;         /* [L] FP32 level scales                */
  const uint32_t *resolution; /* [L]                                  */
  const uint32_t *dense;      /* [L]                                  */
  const uint32_t *size;       /* [L]                                  */
  const uint32_t *offset;     /* [L]                                  */
  const float *table;         /* entries x F (FP16 tables pre-widened) */
  int n_coef;                 /* 6 N^2                                */
  const float *wfeat;         /* [feature_dim][n_coef]                */
  const float *proj;          /* [feature_dim][feature_dim] or NULL   */
  const float *layers;        /* W1 [H][F], W2, W3 [H][H], W4 [3][H]  */
  int feature_dim, hidden;    /* 32, 64                               */
  int n_hidden_layers;        /* 3                                    */
  int softplus;
  int aux;                    /* 1: 13 G-buffer encodings + 3 pads    */
  float aux_pad;              /*    enter the second layer through    */
  const float *waux;          /*    waux [H][16]                      */
} orc_scene;

/* G-buffer arrays (reference layout: float4 except roughness / env_flag). */
typedef struct {
  const float *position, *normal, *reflect, *roughness, *kd, *ks, *direct;
  const uint8_t *env_flag;
} orc_gbuffer;

/* The 13 aux encodings of one pixel plus 3 pads (FP32). */
void orc_aux_inputs(const orc_gbuffer *gb, int p, float pad, float out[16]);

void orc_hashgrid_levels(int n_levels, int log2_hashmap_size, int base_resolution,
                         float per_level_scale, float *scale, uint32_t *resolution,
                         uint32_t *dense, uint32_t *size, uint32_t *offset);

void orc_hash_encode(const orc_scene *sc, const float pos[3], uint32_t *indices,
                     double *features);

void orc_haar_face(const float *src, int n, int stride, float *dst);
void orc_envmap_haar(const float *faces, int n, float *coef);
void orc_rotate_cubemap(const float *src, int n, double yaw, float *dst);
float orc_score(const float *coef, int t, int n);
void orc_topk(const float *coef, int n, int k, int32_t *indices, float *strengths);

void orc_relight_pixels(const orc_scene *sc, const int32_t *sel, const float *light, int k,
                        const orc_gbuffer *gb, const int32_t *pixels, int n_pixels,
                        double *indirect, double *with_direct, double *transport, int n_threads);

uint32_t orc_tonemap(float r, float g, float b);
void orc_tonemap_thresholds(float *out255);

/* Guided a-trous denoiser; images float4 [h][w], out double [h][w][4]. */
void orc_denoise(const float *color, const float *normal, const float *albedo, int w, int h,
                 int iterations, double sigma_color, double sigma_normal, double sigma_albedo,
                 double blend, double *out);

/* Material row: kd[3], ks[3], roughness, type (as a float; 1 = DisneyBSDF,
 * 3 = mirror), specular_transmission, metallic, subsurface, specular,
 * specular_tint, anisotropic, sheen, sheen_tint, clearcoat,
 * clearcoat_gloss, eta (Renderer/Material.h:40-55). */
#define ORC_MAT_FLOATS 19

/* Texture fields of a material (devicePrograms.cu:121-147): kd, ks (RGB),
 * roughness, then the DisneyBSDF parameters and eta (first channel). */
#define ORC_TEX_FIELDS 14

/* Triangle mesh for orc_gbuffer_cast.  materials: [n][ORC_MAT_FLOATS].
 * Optional textures (all four of texcoords, mat_tex, tex_texels and the
 * sizes, or texcoords NULL): texcoords [n_vertices][2]; mat_tex
 * [n_materials][ORC_TEX_FIELDS] texture indices or -1; texture t is
 * tex_w[t] x tex_h[t] RGBA8 texels at tex_texels + tex_offset[t] (bytes),
 * row 0 at v = 0. */
typedef struct {
  const float *vertices;     /* [n_vertices][3]        */
  const float *normals;      /* [n_vertices][3] or NULL */
  const int32_t *indices;    /* [n_tris][3]            */
  int64_t n_tris;
  const int32_t *material_ids; /* [n_tris] or NULL     */
  const float *materials;
  const float *texcoords;
  const int32_t *mat_tex;
  const uint8_t *tex_texels;
  const int64_t *tex_offset;
  const int32_t *tex_w, *tex_h;
} orc_mesh;

/* tex2D of texture t at (u, v) as the reference's CUDA texture objects read
 * 8-bit textures (Renderer/textures.cpp:8-47): wrap, bilinear between texel
 * centres, b / 255 with the sRGB decode on RGB.  out[4]. */
void orc_tex_sample(const orc_mesh *m, int t, double u, double v, double *out);

/* First-hit G-buffer of a w x h view, FP64 brute force (one ray per pixel,
 * jitter at its mean).  camera: from[3], at[3], up[3]; bounds: lower[3],
 * upper[3].  Outputs double [P][4] (position, normal, reflect, kd, ks),
 * double [P] roughness, uint8 [P] env_flag, int32 [P] first-hit triangle
 * (-1: none) and double [P] margin: the smallest barycentric edge distance,
 * relative depth gap or grazing figure met along the pixel's path (small
 * values mark pixels a rounding difference can flip). */
void orc_gbuffer_cast(const orc_mesh *m, const float *camera, int w, int h, int max_bounce,
                      const float *bounds, const float *env_faces, int env_n, double *position,
                      double *normal, double *reflect, double *kd, double *ks, double *roughness,
                      uint8_t *env_flag, int32_t *prim, double *margin);

/* The same plus the direct term (FP64, replaying the product's random
 * stream; light faces [6][light_n][light_n][3], `samples` estimates per
 * pixel); direct: double [P][4]. */
void orc_gbuffer_cast_direct(const orc_mesh *m, const float *camera, int w, int h,
                             int max_bounce, const float *bounds, const float *env_faces,
                             int env_n, const float *light, int light_n, int samples,
                             uint32_t seed, double *position, double *normal, double *reflect,
                             double *kd, double *ks, double *roughness, uint8_t *env_flag,
                             int32_t *prim, double *margin, double *direct);
/* Environment sampling tables (layout: face pmf[6], face cdf[7], face
 * luminance[6], then per face cdf_rows[n][n+1], pdf_rows[n][n],
 * cdf_marginals[n+1], pdf_marginals[n]). */
size_t orc_env_table_doubles(int n);
void orc_env_tables(const float *faces, int n, double *tab);
/* Pieces of the direct term for tests: BSDF eval (with the cosine) and pdf
 * of the material row's type (0 roughPlastic, 2 cookTorrance), and the
 * environment pdf per unit face area. */
void orc_ct_eval(const float *mat, const double *ns, const double *ng, const double *l,
                 const double *v, double *out);
double orc_ct_pdf(const float *mat, const double *ns, const double *ng, const double *l,
                  const double *v);
double orc_env_pdf(const double *tab, int n, const double *d);
/* The product's counter-based uniform in [0, 1). */
float orc_rng_uniform(uint32_t seed, uint32_t pixel, uint32_t counter);

/* pbrt's real SH basis of degree 8 at unit w: 81 values, index l (l+1)+m. */
void orc_sh_eval(const double *w, double *y);
/* SH-PRT render (FP64, replaying the product's stream).  The mesh needs
 * normals.  transfer [n_vertices][nc][nc][3], light_sh [nc][3],
 * nc = (degree+1)^2; color double [P][4], prim int32 [P], margin [P]. */
void orc_sh_render(const orc_mesh *m, const float *camera, int w, int h, const float *transfer,
                   const float *light_sh, int degree, int samples, uint32_t seed,
                   const float *env_faces, int env_n, double *color, int32_t *prim,
                   double *margin);

/* SH transfer matrices (RenderEngine::shComputeTMatrix with
 * __raygen__ComputeTMatrix), FP64, replaying the product's random stream.
 * basis [nc][6][n][n][3] (nc = (degree+1)^2 basis maps); verts: the vertex
 * indices to compute; T double [n_verts][nc * nc][3] (row k * nc + i: SH
 * coefficient k of the start direction, basis map i); abs_sum (or NULL):
 * the same sums of absolute terms; margin (or NULL) double [n_verts][nc]:
 * the smallest robustness figure met by the vertex's paths under map i.
 * rr_depth: Russian roulette after this depth (3 in the reference). */
void orc_sh_transfer(const orc_mesh *m, const float *basis, int n, int degree, int samples,
                     int max_bounce, int rr_depth, uint32_t seed, const int64_t *verts,
                     int64_t n_verts, double *T, double *abs_sum, double *margin);

/* BSDF taps over many directions (mat: a material row; u [n][3] uniforms;
 * l, out [n][3]). */
void orc_bsdf_sample_many(const float *mat, const double *ns, const double *ng, const double *v,
                          int64_t n, const float *u, double *out, int32_t *ok);
void orc_bsdf_eval_many(const float *mat, const double *ns, const double *ng, const double *v,
                        int64_t n, const double *l, double *out);
void orc_bsdf_pdf_many(const float *mat, const double *ns, const double *ng, const double *v,
                       int64_t n, const double *l, double *out);

/* FP32 restatement timed as the CPU baseline (cpu_baseline.c). */
double orc_cpu_relight_fp32(const orc_scene *sc, const int32_t *sel, const float *light, int k,
                            const orc_gbuffer *gb, int pixel_begin, int n_pixels,
                            float *indirect, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
