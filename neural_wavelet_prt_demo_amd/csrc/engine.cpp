// Context, scene / environment-map registries and the C ABI of libnwprt.
//
// Replaces the TCNN half of RenderEngine (Renderer/tinycudann/tcnn.cu:40-162,
// Renderer/RenderEngine.cpp:162-175): instead of process-global network and
// matrices, a context owns device-resident scenes (hash grid, wavelet
// feature table, weights), environment maps (K indices + strengths) and
// per-frame-slot work buffers that grow and are reused.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <tuple>

#include "engine_state.hpp"
#include "kernels.hpp"

namespace nwprt {

void bind(nwprt_ctx ctx) {
  require(ctx != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "null context");
  hip_check(hipSetDevice(ctx->device), "hipSetDevice");
}

hipStream_t pick(nwprt_ctx ctx, void *stream) {
  return stream ? (hipStream_t)stream : ctx->stream;
}

Scene &get_scene(nwprt_ctx ctx, int id) {
  auto it = ctx->scenes.find(id);
  require(it != ctx->scenes.end(), NWPRT_ERR_NOT_FOUND, "unknown scene id " + std::to_string(id));
  return *it->second;
}

EnvMap &get_env(nwprt_ctx ctx, int id) {
  auto it = ctx->envs.find(id);
  require(it != ctx->envs.end(), NWPRT_ERR_NOT_FOUND,
          "unknown environment map id " + std::to_string(id));
  return *it->second;
}

void check_aligned(const void *p, size_t a, const char *what) {
  require(((uintptr_t)p % a) == 0, NWPRT_ERR_INVALID_ARGUMENT,
          std::string(what) + " must be " + std::to_string(a) + "-byte aligned");
}

// Element type of the per-pixel features for a precision (the FP8 form
// reads BF16 features: its first layer is a BF16 MFMA).
Elem elem_of(nwprt_precision p) {
  switch (p) {
    case NWPRT_PREC_FP32: return Elem::F32;
    case NWPRT_PREC_FP16: return Elem::F16;
    case NWPRT_PREC_BF16: return Elem::BF16;
    case NWPRT_PREC_FP8: return Elem::BF16;
  }
  fail(NWPRT_ERR_INVALID_ARGUMENT, "unknown precision");
}

void check_frame(const nwprt_gbuffer *gb, int w, int h, bool need_direct) {
  require(gb != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "null G-buffer");
  require(w > 0 && h > 0 && (long long)w * h <= (1ll << 28), NWPRT_ERR_INVALID_ARGUMENT,
          "frame size must be positive and at most 2^28 pixels");
  require(gb->position != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "G-buffer position is required");
  check_aligned(gb->position, 16, "G-buffer position");
  if (need_direct) {
    require(gb->direct != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "G-buffer direct is required");
    check_aligned(gb->direct, 16, "G-buffer direct");
  }
}

void check_outputs(const nwprt_frame_out &out) {
  if (out.indirect) check_aligned(out.indirect, 16, "indirect output");
  if (out.indirect_with_direct) check_aligned(out.indirect_with_direct, 16, "with-direct output");
  if (out.rgba8) check_aligned(out.rgba8, 4, "rgba8 output");
}

// Hash-grid encode of P pixels of `gb` into `features` ([P][32] of elem).
EncodeArgs encode_args(const Scene &sc, const nwprt_gbuffer *gb, int P, Elem elem,
                       void *features) {
  EncodeArgs a{};
  a.n_pixels = P;
  a.position = gb->position;
  a.env_flag = gb->env_flag;
  a.hash = sc.hash.p;
  a.hash_f16 = sc.desc.hash_dtype == NWPRT_DTYPE_F16;
  a.n_features = sc.desc.grid.n_features_per_level;
  a.lv = sc.lv;
  a.projection = sc.desc.has_projection ? sc.proj.as<float>() : nullptr;
  a.features = features;
  a.out_elem = elem;
  a.planar = planar_features(sc, elem);
  return a;
}

const void *scene_tiles(const Scene &sc, nwprt_precision prec) {
  return prec == NWPRT_PREC_BF16 ? sc.tiles_bf16.p : sc.tiles_f16.p;
}

bool planar_features(const Scene &sc, Elem elem) {
  return elem != Elem::F32 && sc.desc.aux_dims == 0 && sc.desc.grid.n_levels == 16 &&
         sc.desc.grid.n_features_per_level == 2;
}

bool fuse_encode(const Scene &sc, nwprt_precision prec) {
  return elem_of(prec) != Elem::F32 && planar_features(sc, elem_of(prec));
}

void set_fused_encode(RelightArgs &a, const Scene &sc, const nwprt_gbuffer *gb) {
  check_aligned(gb->position, 16, "G-buffer position");
  a.features = nullptr;
  a.features_planar = 1;
  a.fused_encode = 1;
  a.position = gb->position;
  a.hash = sc.hash.p;
  a.hash_f16 = sc.desc.hash_dtype == NWPRT_DTYPE_F16;
  a.lv = sc.lv;
}

const float *scene_projection(const Scene &sc) {
  return sc.desc.has_projection ? sc.proj.as<float>() : nullptr;
}

void set_aux_encode(EncodeArgs &e, const Scene &sc, const nwprt_gbuffer *gb, void *aux) {
  const float *f4[] = {gb->normal, gb->reflect, gb->kd, gb->ks};
  for (const float *p : f4) {
    require(p != nullptr, NWPRT_ERR_INVALID_ARGUMENT,
            "G-buffer normal, reflect, kd and ks are required by this scene's aux inputs");
    check_aligned(p, 16, "G-buffer aux input");
  }
  require(gb->roughness != nullptr, NWPRT_ERR_INVALID_ARGUMENT,
          "G-buffer roughness is required by this scene's aux inputs");
  e.normal = gb->normal;
  e.reflect = gb->reflect;
  e.roughness = gb->roughness;
  e.kd = gb->kd;
  e.ks = gb->ks;
  e.aux = aux;
  e.aux_pad = sc.desc.aux_pad_value;
}

void set_aux_relight(RelightArgs &a, const Scene &sc, nwprt_precision prec, const void *aux) {
  a.aux = aux;
  a.waux_f32 = sc.waux_f32();
  a.waux_tiles = prec == NWPRT_PREC_FP16 ? sc.aux_f16.as<uint16_t>() : sc.aux_bf16.as<uint16_t>();
  a.waux32_tiles = sc.aux32_bf16.as<uint16_t>();
}

size_t fp8_mk_bytes(int k) { return (size_t)k * 4 * 64 * 16; }

size_t fp8_calib_bytes(const Scene &sc, int n_pixels) {
  return (size_t)fp8_calib_count(n_pixels) * (64 + (sc.desc.aux_dims ? 32 : 0));
}

void prepare_fp8(const Scene &sc, const int32_t *idx, int k, const RelightArgs &a,
                 const Fp8Work &w, hipStream_t s) {
  const int P = a.n_pixels;
  const bool planar = planar_features(sc, Elem::BF16);
  launch_prepare_selection(sc.w_f32.as<float>(), sc.wfeat.as<float>(),
                           6 * sc.desc.face_size * sc.desc.face_size, idx, k, Elem::BF16, w.mk16,
                           nullptr, s, w.mk32, planar, scene_projection(sc));
  // Per-layer activation maxima over about 16K pixels spread over the frame.
  const int stride = fp8_calib_stride(P), n_cal = fp8_calib_count(P);
  uint16_t *calib_aux = sc.desc.aux_dims ? w.calib + (size_t)n_cal * 32 : nullptr;
  if (a.fused_encode)
    launch_calib_encode(a, stride, w.calib, n_cal, w.amax, s);
  else
    launch_calib_gather((const uint16_t *)a.features, (const uint16_t *)a.aux, a.env_flag, P,
                        stride, w.calib, calib_aux, n_cal, w.amax, planar, s);
  RelightArgs c{};
  c.n_pixels = n_cal;
  c.k = k;
  c.features = w.calib;
  c.features_planar = planar;
  c.mk = w.mk16;
  c.w_tiles = sc.tiles_bf16.as<uint16_t>();
  c.amax = w.amax;
  if (sc.desc.aux_dims) {
    c.aux = calib_aux;
    c.waux_tiles = sc.aux_bf16.as<uint16_t>();
  }
  launch_calibrate_amax(c, s);
}

void set_fp8_relight(RelightArgs &a, const Scene &sc, const Fp8Work &w) {
  a.mk = w.mk32;
  a.w8_tiles = sc.tiles_fp8.as<uint8_t>();
  for (int i = 0; i < 3; ++i) a.w8_exp[i] = sc.fp8_exp[i];
  a.amax = w.amax;
}

namespace {

const char *status_name(nwprt_status s) {
  switch (s) {
    case NWPRT_OK: return "ok";
    case NWPRT_ERR_INVALID_ARGUMENT: return "invalid argument";
    case NWPRT_ERR_IO: return "i/o error";
    case NWPRT_ERR_FORMAT: return "format error";
    case NWPRT_ERR_DEVICE: return "device error";
    case NWPRT_ERR_OUT_OF_MEMORY: return "out of memory";
    case NWPRT_ERR_UNSUPPORTED: return "unsupported";
    case NWPRT_ERR_STATE: return "invalid call order";
    case NWPRT_ERR_NOT_FOUND: return "unknown handle";
  }
  return "unknown status";
}

// Tuning-table key of a precision: the element type, 3 for the FP8 form.
int tune_class(nwprt_precision p) { return p == NWPRT_PREC_FP8 ? 3 : (int)elem_of(p); }

void create_scene(nwprt_ctx ctx, const nwprt_scene_desc &d, const nwprt_scene_params &prm,
                  int *scene_id) {
  validate_scene_desc(d);
  require(prm.hash_table && prm.wavelet_features && prm.layers && scene_id,
          NWPRT_ERR_INVALID_ARGUMENT, "scene: missing parameter array");
  require(!d.has_projection || prm.projection, NWPRT_ERR_INVALID_ARGUMENT,
          "scene: has_projection set but projection is null");
  auto s = std::make_unique<Scene>();
  s->desc = d;
  hashgrid_levels(d.grid, s->lv);
  const size_t entries = s->lv.offset[s->lv.n_levels];
  const size_t esz = d.hash_dtype == NWPRT_DTYPE_F16 ? 2 : 4;
  s->hash.upload(prm.hash_table, entries * d.grid.n_features_per_level * esz);
  s->wfeat.upload(prm.wavelet_features,
                  (size_t)d.feature_dim * 6 * d.face_size * d.face_size * sizeof(float));
  if (d.has_projection) s->proj.upload(prm.projection, 32 * 32 * sizeof(float));
  const size_t nw = scene_layer_weight_count(d);
  s->w_f32.upload(prm.layers, nw * sizeof(float));
  std::vector<uint16_t> tiles;
  build_weight_tiles(prm.layers, Elem::BF16, tiles);
  s->tiles_bf16.upload(tiles.data(), tiles.size() * 2);
  build_weight_tiles(prm.layers, Elem::F16, tiles);
  s->tiles_f16.upload(tiles.data(), tiles.size() * 2);
  if (d.aux_dims) {
    const float *waux = prm.layers + (nw - 64 * 16);
    build_aux_tiles(waux, Elem::BF16, tiles);
    s->aux_bf16.upload(tiles.data(), tiles.size() * 2);
    build_aux_tiles(waux, Elem::F16, tiles);
    s->aux_f16.upload(tiles.data(), tiles.size() * 2);
    build_fp8_aux_tiles(waux, tiles);
    s->aux32_bf16.upload(tiles.data(), tiles.size() * 2);
  }
  std::vector<uint8_t> t8;
  build_fp8_weight_tiles(prm.layers, t8, s->fp8_exp);
  s->tiles_fp8.upload(t8.data(), t8.size());
  const int id = ctx->next_id++;
  ctx->scenes[id] = std::move(s);
  *scene_id = id;
}

EnvMap &new_env(nwprt_ctx ctx, int n, int k, int *env_id) {
  auto e = std::make_unique<EnvMap>();
  e->n = n;
  e->k = k;
  e->idx.ensure((size_t)k * 4);
  e->strength.ensure((size_t)k * 12);
  hip_check(hipEventCreateWithFlags(&e->ready, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventCreateWithFlags(&e->last_use, hipEventDisableTiming), "hipEventCreate");
  const int id = ctx->next_id++;
  EnvMap &ref = *e;
  ctx->envs[id] = std::move(e);
  *env_id = id;
  return ref;
}

void env_from_faces(nwprt_ctx ctx, const float *faces, nwprt_location loc, int n, int k,
                    int *env_id) {
  require(faces && env_id, NWPRT_ERR_INVALID_ARGUMENT, "envmap: null argument");
  require(n >= 2 && n <= NWPRT_MAX_FACE_SIZE && (n & (n - 1)) == 0, NWPRT_ERR_INVALID_ARGUMENT,
          "envmap: face size must be a power of two in [2, 128]");
  require(k >= 1 && k <= NWPRT_MAX_K && k <= 6 * n * n, NWPRT_ERR_INVALID_ARGUMENT,
          "envmap: K must be in [1, min(1024, 6 N^2)]");
  const size_t m = 6ull * n * n;
  const float *dfaces = faces;
  if (loc == NWPRT_HOST) {
    ctx->staging.ensure(m * 12);
    hip_check(hipMemcpyAsync(ctx->staging.p, faces, m * 12, hipMemcpyHostToDevice, ctx->stream),
              "hipMemcpyAsync faces");
    dfaces = ctx->staging.as<float>();
  } else {
    require(loc == NWPRT_DEVICE, NWPRT_ERR_INVALID_ARGUMENT, "envmap: bad faces location");
  }
  int id = 0;
  EnvMap &e = new_env(ctx, n, k, &id);
  try {
    e.coef.ensure(m * 12);
    e.scores.ensure(envmap_work_bytes(n, 1));
    e.has_coef = true;
    launch_haar(dfaces, m * 3, e.coef.as<float>(), n, 1, e.scores.p, ctx->stream);
    launch_topk(e.coef.as<float>(), n, k, e.scores.p, e.idx.as<int32_t>(),
                e.strength.as<float>(), ctx->stream);
    hip_check(hipEventRecord(e.ready, ctx->stream), "hipEventRecord");
    if (loc == NWPRT_HOST)  // the staging buffer is reused by the next call
      hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  } catch (...) {
    ctx->envs.erase(id);
    throw;
  }
  *env_id = id;
}

void env_from_selection(nwprt_ctx ctx, const int32_t *idx, const float *str, int k, int n,
                        int *env_id) {
  require(idx && str && env_id, NWPRT_ERR_INVALID_ARGUMENT, "envmap: null argument");
  require(n >= 2 && n <= NWPRT_MAX_FACE_SIZE && (n & (n - 1)) == 0, NWPRT_ERR_INVALID_ARGUMENT,
          "envmap: face size must be a power of two in [2, 128]");
  require(k >= 1 && k <= NWPRT_MAX_K, NWPRT_ERR_INVALID_ARGUMENT, "envmap: K must be in [1, 1024]");
  for (int i = 0; i < k; ++i)
    require(idx[i] >= 0 && idx[i] < 6 * n * n, NWPRT_ERR_INVALID_ARGUMENT,
            "envmap: wavelet index " + std::to_string(idx[i]) + " outside [0, 6 N^2)");
  int id = 0;
  EnvMap &e = new_env(ctx, n, k, &id);
  try {
    hip_check(hipMemcpy(e.idx.p, idx, (size_t)k * 4, hipMemcpyHostToDevice), "hipMemcpy idx");
    hip_check(hipMemcpy(e.strength.p, str, (size_t)k * 12, hipMemcpyHostToDevice),
              "hipMemcpy strength");
    hip_check(hipEventRecord(e.ready, ctx->stream), "hipEventRecord");
  } catch (...) {
    ctx->envs.erase(id);
    throw;
  }
  *env_id = id;
}

std::tuple<int, int, int> tune_key(nwprt_precision prec, int k, int n_pixels) {
  int kb = 1, pb = 0;
  while (kb < k) kb <<= 1;
  while ((1ll << pb) < n_pixels) ++pb;
  return {tune_class(prec), kb, pb};
}

// Kernel variant for a relight: the forced one (BF16 / FP16 only), else the
// autotuner's choice for this shape bucket, else 0.
int pick_variant(nwprt_ctx ctx, nwprt_precision prec, int k, int n_pixels) {
  if (ctx->forced_variant >= 0 &&
      (prec != NWPRT_PREC_FP8 || ctx->forced_variant < relight_fp8_variant_count()))
    return ctx->forced_variant;
  auto it = ctx->tuned.find(tune_key(prec, k, n_pixels));
  if (it != ctx->tuned.end()) return it->second;
  return prec == NWPRT_PREC_FP8 ? kDefaultFp8Variant : 0;
}

int variant_count(nwprt_precision prec) {
  return prec == NWPRT_PREC_FP8 ? relight_fp8_variant_count() : relight_mfma_variant_count();
}

void encode_into(nwprt_ctx ctx, Scene &sc, int scene_id, FrameSlot &slot, const nwprt_gbuffer *gb,
                 int w, int h, Elem elem, hipStream_t s, uint32_t *dbg_idx = nullptr,
                 float *dbg_feat = nullptr) {
  const int P = w * h;
  slot.valid = false;
  slot.features.ensure((size_t)P * 32 * (elem == Elem::F32 ? 4 : 2));
  EncodeArgs a = encode_args(sc, gb, P, elem, slot.features.p);
  a.dbg_indices = dbg_idx;
  a.dbg_features = dbg_feat;
  if (sc.desc.aux_dims && dbg_idx == nullptr) {
    slot.aux.ensure((size_t)P * 16 * (elem == Elem::F32 ? 4 : 2));
    set_aux_encode(a, sc, gb, slot.aux.p);
  }
  launch_encode(a, s);
  slot.valid = true;
  slot.scene = scene_id;
  slot.n_pixels = P;
  slot.elem = elem;
}

}  // namespace

void relight_with(nwprt_ctx ctx, const Scene &sc, EnvMap &env, const nwprt_gbuffer *gb, int P,
                  const nwprt_frame_out *out, nwprt_precision prec, nwprt_tonemap_source tm,
                  hipStream_t s, bool fused, const RelightWork &w, float *dump, int dump_begin,
                  int dump_count, int variant) {
  const Elem elem = elem_of(prec);
  require(env.n == sc.desc.face_size, NWPRT_ERR_INVALID_ARGUMENT,
          "environment map face size " + std::to_string(env.n) +
              " does not match the scene's wavelet feature table (" +
              std::to_string(sc.desc.face_size) + ")");
  require(tm == NWPRT_TONEMAP_FULL || tm == NWPRT_TONEMAP_INDIRECT || tm == NWPRT_TONEMAP_NONE,
          NWPRT_ERR_INVALID_ARGUMENT, "unknown tonemap source");
  if (out) check_outputs(*out);
  hip_check(hipStreamWaitEvent(s, env.ready, 0), "hipStreamWaitEvent");
  const int n_coef = 6 * env.n * env.n;
  const bool planar = planar_features(sc, elem);
  RelightArgs a{};
  a.n_pixels = P;
  a.k = env.k;
  a.features = fused ? nullptr : w.features;
  a.features_planar = planar;
  a.light = env.strength.as<float>();
  a.env_flag = gb->env_flag;
  if (fused) {
    require(fuse_encode(sc, prec), NWPRT_ERR_INVALID_ARGUMENT,
            "relight: fused encode needs an MFMA form and a 16 x 2 grid without aux inputs");
    set_fused_encode(a, sc, gb);
  }
  a.direct = gb->direct;
  a.indirect = out ? out->indirect : nullptr;
  a.with_direct = out ? out->indirect_with_direct : nullptr;
  a.rgba8 = out ? out->rgba8 : nullptr;
  a.tonemap_source = (int)tm;
  a.softplus = sc.desc.output_softplus;
  a.dump = dump;
  a.dump_begin = dump_begin;
  a.dump_count = dump_count;
  if (sc.desc.aux_dims) set_aux_relight(a, sc, prec, w.aux);
  if (elem == Elem::F32) {
    launch_prepare_selection(sc.w_f32.as<float>(), sc.wfeat.as<float>(), n_coef,
                             env.idx.as<int32_t>(), env.k, Elem::BF16, nullptr, w.wsel, s);
    a.wsel = w.wsel;
    a.w_f32 = sc.w_f32.as<float>();
    launch_relight_plain(a, s);
  } else if (prec == NWPRT_PREC_FP8) {
    Fp8Work f;
    f.mk16 = w.mk;
    f.mk32 = w.mk32;
    f.calib = w.calib;
    f.amax = w.amax;
    prepare_fp8(sc, env.idx.as<int32_t>(), env.k, a, f, s);
    set_fp8_relight(a, sc, f);
    launch_relight_fp8(a, variant >= 0 ? variant : pick_variant(ctx, prec, env.k, P), s);
  } else {
    launch_prepare_selection(sc.w_f32.as<float>(), sc.wfeat.as<float>(), n_coef,
                             env.idx.as<int32_t>(), env.k, elem, w.mk, nullptr, s, nullptr, planar,
                             scene_projection(sc));
    a.mk = w.mk;
    a.w_tiles = elem == Elem::BF16 ? sc.tiles_bf16.as<uint16_t>() : sc.tiles_f16.as<uint16_t>();
    launch_relight_mfma(a, elem, variant >= 0 ? variant : pick_variant(ctx, prec, env.k, P), s);
  }
  hip_check(hipEventRecord(env.last_use, s), "hipEventRecord");
}

namespace {

// A relight from a frame slot: the slot's features (p1) unless fused, and
// the slot's work buffers, grown as needed.
void relight_from_slot(nwprt_ctx ctx, Scene &sc, int scene_id, EnvMap &env, FrameSlot &slot,
                       const nwprt_gbuffer *gb, int w, int h, const nwprt_frame_out *out,
                       nwprt_precision prec, nwprt_tonemap_source tm, hipStream_t s,
                       float *dump = nullptr, int dump_begin = 0, int dump_count = 0,
                       int variant = -1, bool fused = false) {
  const Elem elem = elem_of(prec);
  const int P = w * h;
  require(fused || (slot.valid && slot.scene == scene_id && slot.n_pixels == P &&
                    slot.elem == elem),
          NWPRT_ERR_STATE,
          "pass p2: frame slot was not encoded (p1) for this scene, size and precision");
  RelightWork wk;
  wk.features = fused ? nullptr : slot.features.p;
  wk.aux = sc.desc.aux_dims ? slot.aux.p : nullptr;
  if (elem == Elem::F32) {
    slot.wsel.ensure((size_t)env.k * 32 * 4);
    wk.wsel = slot.wsel.as<float>();
  } else {
    slot.mk.ensure((size_t)env.k * 4 * 64 * 16);
    wk.mk = slot.mk.as<uint16_t>();
    if (prec == NWPRT_PREC_FP8) {
      slot.mk32.ensure(fp8_mk_bytes(env.k));
      slot.calib.ensure(fp8_calib_bytes(sc, P));
      slot.amax.ensure(16);
      wk.mk32 = slot.mk32.as<uint16_t>();
      wk.calib = slot.calib.as<uint16_t>();
      wk.amax = slot.amax.as<float>();
    }
  }
  relight_with(ctx, sc, env, gb, P, out, prec, tm, s, fused, wk, dump, dump_begin, dump_count,
               variant);
}

// ------------------------------------------------------------------ batch --

constexpr int kDefaultBatchVariant = 5;  // pipe.w4reg.occ2

Batch &get_batch(nwprt_ctx ctx, int id) {
  auto it = ctx->batches.find(id);
  require(it != ctx->batches.end(), NWPRT_ERR_NOT_FOUND, "unknown batch id " + std::to_string(id));
  return *it->second;
}

// Every job's scene and environment map still exist, with the buffers the
// graph captured; the launch stream waits for the maps.
void check_batch_inputs(nwprt_ctx ctx, const Batch &b, hipStream_t s) {
  for (size_t j = 0; j < b.jobs.size(); ++j) {
    const Scene &sc = get_scene(ctx, b.jobs[j].scene_id);
    EnvMap &env = get_env(ctx, b.jobs[j].env_id);
    const void *tiles = b.prec == NWPRT_PREC_FP8 ? sc.tiles_fp8.p : scene_tiles(sc, b.prec);
    require(tiles == b.tiles_at[j] && env.idx.p == b.idx_at[j] && env.k == b.k[j],
            NWPRT_ERR_STATE, "batch: a job's scene or environment map changed since the batch "
                             "was created; create the batch again");
    if (s) hip_check(hipStreamWaitEvent(s, env.ready, 0), "hipStreamWaitEvent");
  }
}

// Work buffers of job j of a batch (FP8 form).
Fp8Work batch_fp8_work(const Batch &b, size_t j) {
  Fp8Work w;
  w.mk16 = (uint16_t *)(b.mk.as<char>() + b.mk_off[j]);
  w.mk32 = (uint16_t *)(b.mk32.as<char>() + b.mk32_off[j]);
  w.calib = (uint16_t *)(b.calib.as<char>() + b.calib_off[j]);
  w.amax = (float *)(b.amax.as<char>() + b.amax_off[j]);
  return w;
}

// Records the batch graph on the context stream: per job the standalone
// encode (jobs the persistent kernel does not encode itself) and the
// selection tiles -- for the FP8 form also the calibration sample and the
// activation maxima -- then the counter reset and the persistent kernel.
void capture_batch(nwprt_ctx ctx, Batch &b, int variant) {
  const bool fp8 = b.prec == NWPRT_PREC_FP8;
  require(fp8 ? variant >= 0 && variant < relight_fp8_variant_count()
              : relight_batch_variant_ok(variant),
          NWPRT_ERR_INVALID_ARGUMENT, "batch: tile variant has no persistent form");
  b.drop_graph();
  const hipStream_t s = ctx->stream;
  const Elem elem = elem_of(b.prec);
  const std::vector<RelightArgs> &args = b.host_args;
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
  try {
    for (size_t j = 0; j < b.jobs.size(); ++j) {
      const nwprt_batch_job &jb = b.jobs[j];
      const Scene &sc = get_scene(ctx, jb.scene_id);
      const EnvMap &env = get_env(ctx, jb.env_id);
      if (!b.fused[j]) {
        EncodeArgs e = encode_args(sc, &jb.gbuffer, b.p[j], elem,
                                   b.features.as<char>() + b.feat_off[j]);
        if (b.aux[j]) set_aux_encode(e, sc, &jb.gbuffer, b.auxbuf.as<char>() + b.aux_off[j]);
        launch_encode(e, s);
      }
      if (fp8) {
        prepare_fp8(sc, env.idx.as<int32_t>(), b.k[j], args[j], batch_fp8_work(b, j), s);
      } else {
        launch_prepare_selection(sc.w_f32.as<float>(), sc.wfeat.as<float>(), 6 * env.n * env.n,
                                 env.idx.as<int32_t>(), b.k[j], elem,
                                 (uint16_t *)(b.mk.as<char>() + b.mk_off[j]), nullptr, s, nullptr,
                                 planar_features(sc, elem), scene_projection(sc));
      }
    }
    unsigned int *counter = (unsigned int *)(b.meta.as<int>() + 1);
    launch_reset_counter(counter, s);
    if (fp8)
      launch_relight_fp8_batch(b.args.as<RelightArgs>(), b.items.as<BatchItem>(),
                               b.meta.as<int>(), counter, b.softplus, b.has_aux, variant,
                               ctx->n_cu, s);
    else
      launch_relight_batch(b.args.as<RelightArgs>(), b.items.as<BatchItem>(), b.meta.as<int>(),
                           counter, elem, b.softplus, b.has_aux, variant, ctx->n_cu, s);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    throw;
  }
  hip_check(hipStreamEndCapture(s, &b.graph), "hipStreamEndCapture");
  hip_check(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  b.variant = variant;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

void create_batch(nwprt_ctx ctx, const nwprt_batch_job *jobs, int n, nwprt_precision prec,
                  nwprt_tonemap_source tm, int *batch_id) {
  require(jobs && batch_id && n >= 1 && n <= 65536, NWPRT_ERR_INVALID_ARGUMENT,
          "batch: need 1 to 65536 jobs");
  require(prec == NWPRT_PREC_BF16 || prec == NWPRT_PREC_FP16 || prec == NWPRT_PREC_FP8,
          NWPRT_ERR_UNSUPPORTED, "batch: BF16, FP16 or FP8 precision only");
  require(tm == NWPRT_TONEMAP_FULL || tm == NWPRT_TONEMAP_INDIRECT || tm == NWPRT_TONEMAP_NONE,
          NWPRT_ERR_INVALID_ARGUMENT, "unknown tonemap source");
  const bool fp8 = prec == NWPRT_PREC_FP8;
  const Elem elem = elem_of(prec);
  auto b = std::make_unique<Batch>();
  b->prec = prec;
  b->tm = tm;
  size_t fo = 0, ao = 0, mo = 0, m32 = 0, co = 0, am = 0;
  for (int j = 0; j < n; ++j) {
    const nwprt_batch_job &jb = jobs[j];
    const Scene &sc = get_scene(ctx, jb.scene_id);
    EnvMap &env = get_env(ctx, jb.env_id);
    require(env.n == sc.desc.face_size, NWPRT_ERR_INVALID_ARGUMENT,
            "batch: environment map face size does not match the job's scene");
    check_frame(&jb.gbuffer, jb.width, jb.height, true);
    check_outputs(jb.out);
    const bool sp = sc.desc.output_softplus != 0;
    if (j == 0) b->softplus = sp;
    require(sp == b->softplus, NWPRT_ERR_UNSUPPORTED,
            "batch: all jobs must share the output activation");
    const int P = jb.width * jb.height;
    const bool fused = fuse_encode(sc, prec), aux = sc.desc.aux_dims != 0;
    if (aux) {  // checks the G-buffer's aux inputs
      EncodeArgs e{};
      set_aux_encode(e, sc, &jb.gbuffer, nullptr);
    }
    b->jobs.push_back(jb);
    b->k.push_back(env.k);
    b->p.push_back(P);
    b->fused.push_back(fused);
    b->aux.push_back(aux);
    b->has_aux = b->has_aux || aux;
    b->feat_off.push_back(fo);
    b->aux_off.push_back(ao);
    b->mk_off.push_back(mo);
    b->mk32_off.push_back(m32);
    b->calib_off.push_back(co);
    b->amax_off.push_back(am);
    b->tiles_at.push_back(fp8 ? sc.tiles_fp8.p : scene_tiles(sc, prec));
    b->idx_at.push_back(env.idx.p);
    if (!fused) fo += align256((size_t)P * 64);
    if (aux) ao += align256((size_t)P * 32);
    mo += align256((size_t)env.k * 4 * 64 * 16);
    if (fp8) {
      m32 += align256(fp8_mk_bytes(env.k));
      co += align256(fp8_calib_bytes(sc, P));
      am += 256;
    }
    b->n_items += (P + 63) / 64;
    b->n_pairs += (long long)P * env.k;
  }
  require(b->n_items < (1ll << 31), NWPRT_ERR_INVALID_ARGUMENT, "batch: too many pixels");
  if (fo) b->features.ensure(fo);
  if (ao) b->auxbuf.ensure(ao);
  b->mk.ensure(mo);
  if (fp8) {
    b->mk32.ensure(m32);
    b->calib.ensure(co);
    b->amax.ensure(am);
  }
  // Device job table.
  b->host_args.assign(n, RelightArgs{});
  for (int j = 0; j < n; ++j) {
    const nwprt_batch_job &jb = b->jobs[j];
    const Scene &sc = get_scene(ctx, jb.scene_id);
    const EnvMap &env = get_env(ctx, jb.env_id);
    RelightArgs &a = b->host_args[j];
    a.n_pixels = b->p[j];
    a.k = b->k[j];
    a.features = b->fused[j] ? nullptr : b->features.as<char>() + b->feat_off[j];
    a.features_planar = planar_features(sc, elem);
    if (b->fused[j]) set_fused_encode(a, sc, &jb.gbuffer);
    if (b->aux[j]) set_aux_relight(a, sc, prec, b->auxbuf.as<char>() + b->aux_off[j]);
    a.mk = (const uint16_t *)(b->mk.as<char>() + b->mk_off[j]);
    a.light = env.strength.as<float>();
    a.w_tiles = (const uint16_t *)scene_tiles(sc, fp8 ? NWPRT_PREC_BF16 : prec);
    a.env_flag = jb.gbuffer.env_flag;
    a.direct = jb.gbuffer.direct;
    a.indirect = jb.out.indirect;
    a.with_direct = jb.out.indirect_with_direct;
    a.rgba8 = jb.out.rgba8;
    a.tonemap_source = (int)tm;
    a.softplus = b->softplus;
  }
  if (fp8) {
    // The preparation reads each job's arguments as they are here (features
    // and fused-encode fields); the job table gets the FP8 fields.
    std::vector<RelightArgs> dev_args = b->host_args;
    for (int j = 0; j < n; ++j)
      set_fp8_relight(dev_args[j], get_scene(ctx, b->jobs[j].scene_id), batch_fp8_work(*b, j));
    b->args.upload(dev_args.data(), dev_args.size() * sizeof(RelightArgs));
  } else {
    b->args.upload(b->host_args.data(), b->host_args.size() * sizeof(RelightArgs));
  }
  // Work items: jobs in decreasing K, so the costliest items are taken
  // first; jobs of equal K grouped by scene (one hash grid in the caches at
  // a time), then in job order; each job's 64-pixel groups in pixel order.
  std::vector<int> order(n);
  for (int j = 0; j < n; ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    if (b->k[x] != b->k[y]) return b->k[x] > b->k[y];
    return b->jobs[x].scene_id < b->jobs[y].scene_id;
  });
  std::vector<BatchItem> items;
  items.reserve((size_t)b->n_items);
  for (int j : order)
    for (int base = 0; base < b->p[j]; base += 64) items.push_back(BatchItem{j, base});
  b->items.upload(items.data(), items.size() * sizeof(BatchItem));
  const int meta[2] = {(int)b->n_items, 0};
  b->meta.upload(meta, sizeof(meta));
  // The graph does not wait on the maps' events; they are complete now, and
  // each launch makes its stream wait for them.
  for (int j = 0; j < n; ++j)
    hip_check(hipEventSynchronize(get_env(ctx, b->jobs[j].env_id).ready), "hipEventSynchronize");
  capture_batch(ctx, *b, fp8 ? kDefaultFp8Variant : kDefaultBatchVariant);
  const int id = ctx->next_id++;
  ctx->batches[id] = std::move(b);
  *batch_id = id;
}

}  // namespace
}  // namespace nwprt

using namespace nwprt;

extern "C" {

int32_t nwprt_abi_version(void) { return NWPRT_ABI_VERSION; }

const char *nwprt_status_string(nwprt_status s) { return status_name(s); }

nwprt_status nwprt_create(int device, nwprt_ctx *out) {
  return guarded([&] {
    require(out != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "null output pointer");
    int count = 0;
    hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    require(device >= 0 && device < count, NWPRT_ERR_INVALID_ARGUMENT,
            "device " + std::to_string(device) + " not present (" + std::to_string(count) +
                " visible)");
    auto ctx = std::make_unique<nwprt_context>();
    ctx->device = device;
    hip_check(hipSetDevice(device), "hipSetDevice");
    scratch_init(device);
    hip_check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, device),
              "hipDeviceGetAttribute");
    *out = ctx.release();
  });
}

nwprt_status nwprt_destroy(nwprt_ctx ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipDeviceSynchronize();
    hipStream_t s = ctx->stream;
    delete ctx;
    (void)hipStreamDestroy(s);
  });
}

nwprt_status nwprt_sync(nwprt_ctx ctx) {
  return guarded([&] {
    bind(ctx);
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  });
}

nwprt_status nwprt_get_stream(nwprt_ctx ctx, void **stream) {
  return guarded([&] {
    require(ctx && stream, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    *stream = (void *)ctx->stream;
  });
}

nwprt_status nwprt_scene_create(nwprt_ctx ctx, const nwprt_scene_desc *desc,
                                const nwprt_scene_params *params, int32_t *scene_id) {
  return guarded([&] {
    bind(ctx);
    require(desc && params, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    int id = 0;
    create_scene(ctx, *desc, *params, &id);
    *scene_id = id;
  });
}

nwprt_status nwprt_scene_load_folder(nwprt_ctx ctx, const char *model_dir, int32_t *scene_id) {
  return guarded([&] {
    bind(ctx);
    require(model_dir && scene_id, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    const HostScene hs = load_model_folder(model_dir);
    nwprt_scene_params p{};
    p.hash_table = hs.hash_table.data();
    p.wavelet_features = hs.wavelet_features.data();
    p.projection = hs.projection.empty() ? nullptr : hs.projection.data();
    p.layers = hs.layers.data();
    int id = 0;
    create_scene(ctx, hs.desc, p, &id);
    *scene_id = id;
  });
}

nwprt_status nwprt_scene_destroy(nwprt_ctx ctx, int32_t scene_id) {
  return guarded([&] {
    bind(ctx);
    require(get_scene(ctx, scene_id).server_refs == 0, NWPRT_ERR_STATE,
            "scene is used by a queued or running server request");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    ctx->scenes.erase(scene_id);
    for (auto &sl : ctx->slots)
      if (sl.scene == scene_id) sl.valid = false;
  });
}

nwprt_status nwprt_scene_get_desc(nwprt_ctx ctx, int32_t scene_id, nwprt_scene_desc *desc) {
  return guarded([&] {
    require(ctx && desc, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    *desc = get_scene(ctx, scene_id).desc;
  });
}

nwprt_status nwprt_envmap_from_faces(nwprt_ctx ctx, const float *faces,
                                     nwprt_location faces_location, int32_t face_size,
                                     int32_t k, int32_t *env_id) {
  return guarded([&] {
    bind(ctx);
    int id = 0;
    env_from_faces(ctx, faces, faces_location, face_size, k, &id);
    *env_id = id;
  });
}

nwprt_status nwprt_envmap_update_faces(nwprt_ctx ctx, int32_t env_id, const float *faces,
                                       nwprt_location faces_location, void *stream) {
  return guarded([&] {
    bind(ctx);
    EnvMap &e = get_env(ctx, env_id);
    require(e.has_coef, NWPRT_ERR_STATE, "environment map was not built from faces");
    require(e.server_refs == 0, NWPRT_ERR_STATE,
            "environment map is used by a queued or running server request");
    require(faces != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "null faces");
    const hipStream_t s = pick(ctx, stream);
    hip_check(hipStreamWaitEvent(s, e.last_use, 0), "hipStreamWaitEvent");
    const size_t m = 6ull * e.n * e.n;
    const float *dfaces = faces;
    if (faces_location == NWPRT_HOST) {
      ctx->staging.ensure(m * 12);
      hip_check(hipMemcpyAsync(ctx->staging.p, faces, m * 12, hipMemcpyHostToDevice, s),
                "hipMemcpyAsync faces");
      dfaces = ctx->staging.as<float>();
    } else {
      require(faces_location == NWPRT_DEVICE, NWPRT_ERR_INVALID_ARGUMENT, "bad faces location");
    }
    launch_haar(dfaces, m * 3, e.coef.as<float>(), e.n, 1, e.scores.p, s);
    launch_topk(e.coef.as<float>(), e.n, e.k, e.scores.p, e.idx.as<int32_t>(),
                e.strength.as<float>(), s);
    hip_check(hipEventRecord(e.ready, s), "hipEventRecord");
    if (faces_location == NWPRT_HOST) hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  });
}

nwprt_status nwprt_envmap_update_rotated(nwprt_ctx ctx, int32_t env_id, const float *faces,
                                         double yaw, void *stream) {
  return guarded([&] {
    bind(ctx);
    EnvMap &e = get_env(ctx, env_id);
    require(e.has_coef, NWPRT_ERR_STATE, "environment map was not built from faces");
    require(e.server_refs == 0, NWPRT_ERR_STATE,
            "environment map is used by a queued or running server request");
    require(faces != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "null faces");
    require(std::isfinite(yaw), NWPRT_ERR_INVALID_ARGUMENT, "yaw must be finite");
    const hipStream_t s = pick(ctx, stream);
    hip_check(hipStreamWaitEvent(s, e.last_use, 0), "hipStreamWaitEvent");
    const size_t m = 6ull * e.n * e.n;
    e.rotated.ensure(m * 12);
    launch_rotate_cubemap(faces, e.rotated.as<float>(), e.n, yaw, s);
    launch_haar(e.rotated.as<float>(), m * 3, e.coef.as<float>(), e.n, 1, e.scores.p, s);
    launch_topk(e.coef.as<float>(), e.n, e.k, e.scores.p, e.idx.as<int32_t>(),
                e.strength.as<float>(), s);
    hip_check(hipEventRecord(e.ready, s), "hipEventRecord");
  });
}

nwprt_status nwprt_envmap_from_selection(nwprt_ctx ctx, const int32_t *indices,
                                         const float *strengths, int32_t k, int32_t face_size,
                                         int32_t *env_id) {
  return guarded([&] {
    bind(ctx);
    int id = 0;
    env_from_selection(ctx, indices, strengths, k, face_size, &id);
    *env_id = id;
  });
}

nwprt_status nwprt_envmap_load_folder(nwprt_ctx ctx, const char *folder, const char *name,
                                      int32_t k, int32_t *env_id) {
  return guarded([&] {
    bind(ctx);
    require(folder && name && env_id, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    const std::string base(folder), nm(name);
    const HdrImage cross = hdr_read(base + "/cube/" + nm + ".hdr");
    const int n = cross.height / 3;
    std::vector<float> faces((size_t)6 * n * n * 3);
    cross_to_faces(cross.rgb.data(), cross.width, cross.height, faces.data());
    const std::string cpath = base + "/waveletcoeffs/" + nm + ".npy";
    const std::string spath = base + "/waveletstrength/" + nm + ".npy";
    std::ifstream c1(cpath), c2(spath);
    int id = 0;
    if (c1.good() && c2.good()) {
      const NpyArray ci = npy_read(cpath), si = npy_read(spath);
      require(si.shape.size() == 2 && si.shape[1] == 3, NWPRT_ERR_FORMAT,
              "waveletstrength: expected shape [K, 3]");
      const std::vector<int32_t> idx = ci.as_i32();
      const std::vector<float> str = si.as_f32();
      require(idx.size() * 3 == str.size(), NWPRT_ERR_FORMAT,
              "waveletcoeffs / waveletstrength: K mismatch");
      env_from_selection(ctx, idx.data(), str.data(), (int)idx.size(), n, &id);
    } else {
      env_from_faces(ctx, faces.data(), NWPRT_HOST, n, k, &id);
    }
    *env_id = id;
  });
}

nwprt_status nwprt_envmap_haar(nwprt_ctx ctx, const float *faces, int32_t face_size,
                               int32_t n_maps, float *coefficients, void *stream) {
  return guarded([&] {
    bind(ctx);
    require(faces && coefficients, NWPRT_ERR_INVALID_ARGUMENT, "haar: null buffer");
    require(face_size >= 2 && face_size <= NWPRT_MAX_FACE_SIZE && (face_size & (face_size - 1)) == 0,
            NWPRT_ERR_INVALID_ARGUMENT, "haar: face size must be a power of two in [2, 128]");
    require(n_maps >= 1 && n_maps <= 65535, NWPRT_ERR_INVALID_ARGUMENT,
            "haar: n_maps must be in [1, 65535]");
    const size_t m = 6ull * face_size * face_size;
    const char *f = (const char *)faces, *c = (const char *)coefficients;
    const size_t bytes = m * 12 * n_maps;
    require(c + bytes <= f || f + bytes <= c, NWPRT_ERR_INVALID_ARGUMENT,
            "haar: coefficients overlap the faces");
    ctx->haar_work.ensure(envmap_work_bytes(face_size, n_maps));
    launch_haar(faces, m * 3, coefficients, face_size, n_maps, ctx->haar_work.p,
                pick(ctx, stream));
  });
}

nwprt_status nwprt_envmap_destroy(nwprt_ctx ctx, int32_t env_id) {
  return guarded([&] {
    bind(ctx);
    require(get_env(ctx, env_id).server_refs == 0, NWPRT_ERR_STATE,
            "environment map is used by a queued or running server request");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    ctx->envs.erase(env_id);
  });
}

nwprt_status nwprt_envmap_info(nwprt_ctx ctx, int32_t env_id, int32_t *face_size, int32_t *k) {
  return guarded([&] {
    require(ctx, NWPRT_ERR_INVALID_ARGUMENT, "null context");
    const EnvMap &e = get_env(ctx, env_id);
    if (face_size) *face_size = e.n;
    if (k) *k = e.k;
  });
}

nwprt_status nwprt_envmap_get_selection(nwprt_ctx ctx, int32_t env_id, int32_t *indices,
                                        float *strengths) {
  return guarded([&] {
    bind(ctx);
    EnvMap &e = get_env(ctx, env_id);
    hip_check(hipEventSynchronize(e.ready), "hipEventSynchronize");
    if (indices)
      hip_check(hipMemcpy(indices, e.idx.p, (size_t)e.k * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    if (strengths)
      hip_check(hipMemcpy(strengths, e.strength.p, (size_t)e.k * 12, hipMemcpyDeviceToHost),
                "hipMemcpy");
  });
}

nwprt_status nwprt_envmap_get_coefficients(nwprt_ctx ctx, int32_t env_id, float *coefficients) {
  return guarded([&] {
    bind(ctx);
    EnvMap &e = get_env(ctx, env_id);
    require(e.has_coef, NWPRT_ERR_STATE, "environment map was not built from faces");
    require(coefficients, NWPRT_ERR_INVALID_ARGUMENT, "null output");
    hip_check(hipEventSynchronize(e.ready), "hipEventSynchronize");
    hip_check(hipMemcpy(coefficients, e.coef.p, (size_t)6 * e.n * e.n * 12, hipMemcpyDeviceToHost),
              "hipMemcpy");
  });
}

nwprt_status nwprt_relight(nwprt_ctx ctx, int32_t scene_id, int32_t env_id,
                           const nwprt_gbuffer *gbuffer, int32_t width, int32_t height,
                           const nwprt_frame_out *out, nwprt_precision precision,
                           nwprt_tonemap_source tonemap, void *stream) {
  return guarded([&] {
    bind(ctx);
    Scene &sc = get_scene(ctx, scene_id);
    EnvMap &env = get_env(ctx, env_id);
    check_frame(gbuffer, width, height, true);
    const Elem elem = elem_of(precision);
    const hipStream_t s = pick(ctx, stream);
    const bool fused = fuse_encode(sc, precision);
    if (fused)
      ctx->slots[2].valid = false;
    else
      encode_into(ctx, sc, scene_id, ctx->slots[2], gbuffer, width, height, elem, s);
    relight_from_slot(ctx, sc, scene_id, env, ctx->slots[2], gbuffer, width, height, out,
                      precision, tonemap, s, nullptr, 0, 0, -1, fused);
  });
}

nwprt_status nwprt_pass_p1(nwprt_ctx ctx, int32_t scene_id, int32_t frame_slot,
                           const nwprt_gbuffer *gbuffer, int32_t width, int32_t height,
                           nwprt_precision precision, void *stream) {
  return guarded([&] {
    bind(ctx);
    require(frame_slot == 0 || frame_slot == 1, NWPRT_ERR_INVALID_ARGUMENT,
            "frame slot must be 0 or 1");
    Scene &sc = get_scene(ctx, scene_id);
    check_frame(gbuffer, width, height, false);
    encode_into(ctx, sc, scene_id, ctx->slots[frame_slot], gbuffer, width, height,
                elem_of(precision), pick(ctx, stream));
  });
}

nwprt_status nwprt_pass_p2(nwprt_ctx ctx, int32_t scene_id, int32_t env_id, int32_t frame_slot,
                           const nwprt_gbuffer *gbuffer, int32_t width, int32_t height,
                           const nwprt_frame_out *out, nwprt_precision precision,
                           nwprt_tonemap_source tonemap, void *stream) {
  return guarded([&] {
    bind(ctx);
    require(frame_slot == 0 || frame_slot == 1, NWPRT_ERR_INVALID_ARGUMENT,
            "frame slot must be 0 or 1");
    Scene &sc = get_scene(ctx, scene_id);
    EnvMap &env = get_env(ctx, env_id);
    check_frame(gbuffer, width, height, true);
    relight_from_slot(ctx, sc, scene_id, env, ctx->slots[frame_slot], gbuffer, width, height, out,
                      precision, tonemap, pick(ctx, stream));
  });
}

int32_t nwprt_relight_variant_count(void) { return relight_mfma_variant_count(); }

const char *nwprt_relight_variant_name(int32_t variant) {
  return relight_mfma_variant_name(variant);
}

int32_t nwprt_relight_fp8_variant_count(void) { return relight_fp8_variant_count(); }

const char *nwprt_relight_fp8_variant_name(int32_t variant) {
  return relight_fp8_variant_name(variant);
}

nwprt_status nwprt_set_relight_variant(nwprt_ctx ctx, int32_t variant) {
  return guarded([&] {
    require(ctx != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "null context");
    require(variant >= -1 && variant < relight_mfma_variant_count(), NWPRT_ERR_INVALID_ARGUMENT,
            "unknown relight variant " + std::to_string(variant));
    ctx->forced_variant = variant;
  });
}

nwprt_status nwprt_selected_relight_variant(nwprt_ctx ctx, nwprt_precision precision, int32_t k,
                                            int32_t n_pixels, int32_t *variant) {
  return guarded([&] {
    require(ctx && variant && k >= 1 && n_pixels >= 1, NWPRT_ERR_INVALID_ARGUMENT,
            "bad argument");
    *variant = pick_variant(ctx, precision, k, n_pixels);
  });
}

nwprt_status nwprt_autotune_relight(nwprt_ctx ctx, int32_t scene_id, int32_t env_id,
                                    const nwprt_gbuffer *gbuffer, int32_t width, int32_t height,
                                    nwprt_precision precision, int32_t *best_variant,
                                    float *best_ms) {
  return guarded([&] {
    bind(ctx);
    Scene &sc = get_scene(ctx, scene_id);
    EnvMap &env = get_env(ctx, env_id);
    check_frame(gbuffer, width, height, true);
    const Elem elem = elem_of(precision);
    require(elem != Elem::F32, NWPRT_ERR_INVALID_ARGUMENT,
            "autotune: the plain FP32 form has no tile variants");
    const hipStream_t s = ctx->stream;
    FrameSlot &slot = ctx->slots[2];
    // Tune the form nwprt_relight runs (with the fused encode where it applies).
    const bool fused = fuse_encode(sc, precision);
    if (fused)
      slot.valid = false;
    else
      encode_into(ctx, sc, scene_id, slot, gbuffer, width, height, elem, s);
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    int best = 0;
    float best_t = 0.f;
    try {
      for (int v = 0; v < variant_count(precision); ++v) {
        relight_from_slot(ctx, sc, scene_id, env, slot, gbuffer, width, height, nullptr,
                          precision, NWPRT_TONEMAP_NONE, s, nullptr, 0, 0, v, fused);
        float t = 0.f;
        for (int rep = 0; rep < 3; ++rep) {
          hip_check(hipEventRecord(e0, s), "hipEventRecord");
          relight_from_slot(ctx, sc, scene_id, env, slot, gbuffer, width, height, nullptr,
                            precision, NWPRT_TONEMAP_NONE, s, nullptr, 0, 0, v, fused);
          hip_check(hipEventRecord(e1, s), "hipEventRecord");
          hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
          float ms = 0.f;
          hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
          t = rep == 0 ? ms : std::min(t, ms);
        }
        if (v == 0 || t < best_t) {
          best = v;
          best_t = t;
        }
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->tuned[tune_key(precision, env.k, width * height)] = best;
    if (best_variant) *best_variant = best;
    if (best_ms) *best_ms = best_t;
  });
}

nwprt_status nwprt_tonemap(nwprt_ctx ctx, const float *src_float4, uint32_t *dst_rgba8,
                           int32_t width, int32_t height, void *stream) {
  return guarded([&] {
    bind(ctx);
    require(src_float4 && dst_rgba8 && width > 0 && height > 0, NWPRT_ERR_INVALID_ARGUMENT,
            "tonemap: bad argument");
    check_aligned(src_float4, 16, "tonemap source");
    launch_tonemap(src_float4, dst_rgba8, width * height, pick(ctx, stream));
  });
}

void nwprt_denoise_default_params(nwprt_denoise_params *params) {
  if (params) *params = nwprt_denoise_params{5, 1.0f, 0.3f, 0.1f};
}

nwprt_status nwprt_denoise(nwprt_ctx ctx, const float *color, const float *normal,
                           const float *albedo, float *output, int32_t width, int32_t height,
                           float blend_factor, const nwprt_denoise_params *params,
                           void *stream) {
  return guarded([&] {
    bind(ctx);
    nwprt_denoise_params pr;
    nwprt_denoise_default_params(&pr);
    if (params) pr = *params;
    require(color && normal && albedo && output, NWPRT_ERR_INVALID_ARGUMENT,
            "denoise: null buffer");
    require(width > 0 && height > 0 && (int64_t)width * height <= (1ll << 31),
            NWPRT_ERR_INVALID_ARGUMENT, "denoise: bad image size");
    require(pr.iterations >= 1 && pr.iterations <= kMaxDenoisePasses,
            NWPRT_ERR_INVALID_ARGUMENT, "denoise: iterations must be 1..8");
    require(pr.sigma_color > 0.f && pr.sigma_normal > 0.f && pr.sigma_albedo > 0.f &&
                std::isfinite(pr.sigma_color) && std::isfinite(pr.sigma_normal) &&
                std::isfinite(pr.sigma_albedo),
            NWPRT_ERR_INVALID_ARGUMENT, "denoise: sigmas must be positive and finite");
    require(blend_factor >= 0.f && blend_factor <= 1.f, NWPRT_ERR_INVALID_ARGUMENT,
            "denoise: blend factor must be in [0, 1]");
    for (const void *p : {(const void *)color, (const void *)normal, (const void *)albedo,
                          (const void *)output})
      check_aligned(p, 16, "denoise image");
    const size_t bytes = (size_t)width * height * 16;
    const char *o = (const char *)output;
    for (const float *in : {color, normal, albedo}) {
      const char *i = (const char *)in;
      require(o + bytes <= i || i + bytes <= o, NWPRT_ERR_INVALID_ARGUMENT,
              "denoise: output overlaps an input");
    }
    ctx->denoise_scratch.ensure(denoise_scratch_bytes(width, height));
    DenoiseArgs a{color, normal, albedo, output, width, height, pr.iterations,
                  pr.sigma_color, pr.sigma_normal, pr.sigma_albedo, blend_factor};
    launch_denoise(a, ctx->denoise_scratch.p, pick(ctx, stream));
  });
}

nwprt_status nwprt_debug_hash_encode(nwprt_ctx ctx, int32_t scene_id, const float *positions,
                                     int32_t n, uint32_t *indices, float *features,
                                     void *stream) {
  return guarded([&] {
    bind(ctx);
    Scene &sc = get_scene(ctx, scene_id);
    require(positions && n > 0 && (indices || features), NWPRT_ERR_INVALID_ARGUMENT,
            "debug_hash_encode: bad argument");
    check_aligned(positions, 16, "positions");
    nwprt_gbuffer gb{};
    gb.position = (float *)positions;
    DevBuf idx_tmp;
    uint32_t *idx = indices;
    if (!idx) {
      idx_tmp.ensure((size_t)n * sc.lv.n_levels * 8 * 4);
      idx = idx_tmp.as<uint32_t>();
    }
    const hipStream_t s = pick(ctx, stream);
    encode_into(ctx, sc, scene_id, ctx->slots[2], &gb, n, 1, Elem::F32, s, idx, features);
    ctx->slots[2].valid = false;
    if (!indices) hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  });
}

nwprt_status nwprt_debug_transport(nwprt_ctx ctx, int32_t scene_id, int32_t env_id,
                                   const nwprt_gbuffer *gbuffer, int32_t width, int32_t height,
                                   int32_t pixel_begin, int32_t n_pixels, float *transport,
                                   nwprt_precision precision, void *stream) {
  return guarded([&] {
    bind(ctx);
    Scene &sc = get_scene(ctx, scene_id);
    EnvMap &env = get_env(ctx, env_id);
    check_frame(gbuffer, width, height, true);
    require(transport && n_pixels > 0 && pixel_begin >= 0 &&
                (long long)pixel_begin + n_pixels <= (long long)width * height,
            NWPRT_ERR_INVALID_ARGUMENT, "debug_transport: pixel range outside the frame");
    const Elem elem = elem_of(precision);
    const hipStream_t s = pick(ctx, stream);
    encode_into(ctx, sc, scene_id, ctx->slots[2], gbuffer, width, height, elem, s);
    relight_from_slot(ctx, sc, scene_id, env, ctx->slots[2], gbuffer, width, height, nullptr,
                      precision, NWPRT_TONEMAP_NONE, s, transport, pixel_begin, n_pixels);
  });
}

nwprt_status nwprt_batch_create(nwprt_ctx ctx, const nwprt_batch_job *jobs, int32_t n_jobs,
                                nwprt_precision precision, nwprt_tonemap_source tonemap,
                                int32_t *batch_id) {
  return guarded([&] {
    bind(ctx);
    int id = 0;
    create_batch(ctx, jobs, n_jobs, precision, tonemap, &id);
    *batch_id = id;
  });
}

nwprt_status nwprt_batch_launch(nwprt_ctx ctx, int32_t batch_id, void *stream) {
  return guarded([&] {
    bind(ctx);
    Batch &b = get_batch(ctx, batch_id);
    const hipStream_t s = pick(ctx, stream);
    check_batch_inputs(ctx, b, s);
    hip_check(hipGraphLaunch(b.exec, s), "hipGraphLaunch");
    for (const nwprt_batch_job &jb : b.jobs)
      hip_check(hipEventRecord(get_env(ctx, jb.env_id).last_use, s), "hipEventRecord");
  });
}

nwprt_status nwprt_batch_autotune(nwprt_ctx ctx, int32_t batch_id, int32_t *best_variant,
                                  float *best_ms) {
  return guarded([&] {
    bind(ctx);
    Batch &b = get_batch(ctx, batch_id);
    const hipStream_t s = ctx->stream;
    check_batch_inputs(ctx, b, s);
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    int best = -1;
    float best_t = 0.f;
    try {
      const bool fp8 = b.prec == NWPRT_PREC_FP8;
      const int nv = fp8 ? relight_fp8_variant_count() : relight_mfma_variant_count();
      for (int v = 0; v < nv; ++v) {
        if (!fp8 && !relight_batch_variant_ok(v)) continue;
        capture_batch(ctx, b, v);
        hip_check(hipGraphLaunch(b.exec, s), "hipGraphLaunch");
        float t = 0.f;
        for (int rep = 0; rep < 3; ++rep) {
          hip_check(hipEventRecord(e0, s), "hipEventRecord");
          hip_check(hipGraphLaunch(b.exec, s), "hipGraphLaunch");
          hip_check(hipEventRecord(e1, s), "hipEventRecord");
          hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
          float ms = 0.f;
          hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
          t = rep == 0 ? ms : std::min(t, ms);
        }
        if (best < 0 || t < best_t) {
          best = v;
          best_t = t;
        }
      }
      capture_batch(ctx, b, best);
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (best_variant) *best_variant = best;
    if (best_ms) *best_ms = best_t;
  });
}

nwprt_status nwprt_batch_info(nwprt_ctx ctx, int32_t batch_id, int64_t *n_items,
                              int64_t *n_pairs, int32_t *variant) {
  return guarded([&] {
    require(ctx != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "null context");
    const Batch &b = get_batch(ctx, batch_id);
    if (n_items) *n_items = b.n_items;
    if (n_pairs) *n_pairs = b.n_pairs;
    if (variant) *variant = b.variant;
  });
}

nwprt_status nwprt_batch_destroy(nwprt_ctx ctx, int32_t batch_id) {
  return guarded([&] {
    bind(ctx);
    get_batch(ctx, batch_id);
    // A launch may still be running on any stream.
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    ctx->batches.erase(batch_id);
  });
}

}  // extern "C"
