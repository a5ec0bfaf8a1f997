// Host-side state of a libnwprt context, shared by the C ABI (engine.cpp)
// and the request server (server.cpp): device buffers, resident scenes and
// environment maps, frame slots, captured batches, and the helpers every
// entry point uses (error guard, handle lookup, argument checks).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "kernels.hpp"

namespace nwprt {

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t n) {
    if (n <= bytes) return;
    release();
    const hipError_t e = hipMalloc(&p, n);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      fail(NWPRT_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(n) + " bytes failed");
    }
    hip_check(e, "hipMalloc");
    bytes = n;
  }
  void upload(const void *src, size_t n) {
    ensure(n);
    hip_check(hipMemcpy(p, src, n, hipMemcpyHostToDevice), "hipMemcpy H2D");
  }
  template <class T>
  T *as() const { return (T *)p; }
};

struct Scene {
  nwprt_scene_desc desc{};
  nwprt_hashgrid_levels lv{};
  DevBuf hash, wfeat, proj, w_f32, tiles_bf16, tiles_f16, aux_bf16, aux_f16;
  // FP8 form: E4M3 tiles of layers 2..4 and their power-of-two exponents;
  // with aux inputs, the aux columns as BF16 32x32x16 tiles.
  DevBuf tiles_fp8, aux32_bf16;
  int fp8_exp[3] = {0, 0, 0};
  // Queued or running server requests that use this scene: destroy refuses
  // while any do.
  int server_refs = 0;
  // FP32 aux columns of the second layer ([64][16]) inside w_f32, or null.
  const float *waux_f32() const {
    return desc.aux_dims ? w_f32.as<float>() + (scene_layer_weight_count(desc) - 64 * 16) : nullptr;
  }
};

struct EnvMap {
  int n = 0, k = 0;
  bool has_coef = false;
  DevBuf idx, strength, coef;
  // Faces-built maps: Haar / top-K work buffer and, once rotated, the rotated faces
  // (per map, so updates of different maps on different streams never share
  // scratch memory).
  DevBuf scores, rotated;
  // ready: the selection is written (readers wait on it); last_use: the last
  // relight or batch launch that reads it was queued (in-place updates wait
  // on it, so a reader on another stream never sees a half-written one).
  hipEvent_t ready = nullptr, last_use = nullptr;
  // Queued or running server requests that read this map: destroy and
  // in-place updates refuse while any do.
  int server_refs = 0;
  ~EnvMap() {
    if (ready) (void)hipEventDestroy(ready);
    if (last_use) (void)hipEventDestroy(last_use);
  }
};

// A triangle mesh resident for the G-buffer producer: its BVH (two-child
// nodes, triangles in leaf order) and the per-vertex / per-triangle data the
// closest-hit step reads.
struct Mesh {
  DevBuf nodes, tris, indices, vertices, normals, tri_material, materials;
  // Distinct (vertex, material) pairs of the triangle corners, for the SH
  // render: corner_key[3 t + c] -> pair; key_vertex / key_material per pair.
  DevBuf corner_key, key_vertex, key_material;
  int64_t n_keys = 0;
  float lower[3] = {0, 0, 0}, upper[3] = {0, 0, 0};
  int64_t n_tris = 0, n_nodes = 0, n_vertices = 0;
  int depth = 0;
  bool has_normals = false, has_tri_material = false;
  bool sh_ok = true;    // no mirror material (the SH render has no mirror BSDF)
  bool disney = false;  // some material the triangles use is DisneyBSDF
  int32_t n_materials = 0;
  // Material textures (nwprt_mesh_set_textures): texcoords [n_vertices][2]
  // and TexRef [n_materials][kTexFields]; the texture ids the mesh holds.
  DevBuf texcoords, mat_tex;
  std::vector<int32_t> texture_ids;
  bool textured = false;
};

// A material texture (nwprt_texture_create): RGBA8 texels, row 0 at v = 0.
struct TextureObj {
  DevBuf rgba;
  int32_t width = 0, height = 0;
  int uses = 0;  // meshes holding it
};

// Work buffers of one frame slot (0, 1: the caller's two-frame pipeline,
// 2: nwprt_relight / debug taps).
struct FrameSlot {
  DevBuf features, aux, mk, wsel;
  // FP8 form: first-layer 32x32 tiles, calibration subsample (features, then
  // aux inputs), activation maxima.
  DevBuf mk32, calib, amax;
  bool valid = false;
  int scene = -1, n_pixels = 0;
  Elem elem = Elem::F32;
};

// A captured batch of relight jobs (nwprt_batch_*): concatenated per-job
// features and selection tiles, the device job table and work-item list,
// and the graph that encodes every job and runs the persistent kernel.
struct Batch {
  nwprt_precision prec = NWPRT_PREC_BF16;
  nwprt_tonemap_source tm = NWPRT_TONEMAP_FULL;
  bool softplus = true;
  std::vector<nwprt_batch_job> jobs;
  std::vector<int> k, p;
  std::vector<char> fused;  // job encodes inside the persistent kernel
  std::vector<char> aux;    // job's scene has aux inputs
  // Per-job offsets in bytes: features, aux inputs, selection tiles, and for
  // the FP8 form the 32x32 tiles, calibration sample and maxima.
  std::vector<size_t> feat_off, aux_off, mk_off, mk32_off, calib_off, amax_off;
  bool has_aux = false;
  // What the graph captured of each job's scene and map: a launch checks
  // that these buffers are still the ones in use.
  std::vector<const void *> tiles_at, idx_at;
  DevBuf features, auxbuf, mk, mk32, calib, amax, args, items,
      meta;  // meta: int n_items | uint counter
  // Host copy of the job table before the FP8 fields (the preparation reads
  // each job's features and fused-encode fields from it).
  std::vector<RelightArgs> host_args;
  long long n_items = 0, n_pairs = 0;
  int variant = 5;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~Batch() { drop_graph(); }
};

// Request server (server.cpp); owned through a deleter defined there.
struct Server;
void destroy_server(Server *s);
using ServerPtr = std::unique_ptr<Server, void (*)(Server *)>;

}  // namespace nwprt

struct nwprt_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::map<int, std::unique_ptr<nwprt::Scene>> scenes;
  std::map<int, std::unique_ptr<nwprt::EnvMap>> envs;
  int next_id = 1;
  nwprt::FrameSlot slots[3];
  nwprt::DevBuf staging;
  // Tile tuning: forced variant (-1: none) and the autotuner's choices keyed
  // by (element type, K rounded up to a power of two, ceil(log2 pixels)).
  int forced_variant = -1;
  std::map<std::tuple<int, int, int>, int> tuned;
  std::map<int, std::unique_ptr<nwprt::Batch>> batches;
  std::map<int, nwprt::ServerPtr> servers;
  int n_cu = 0;  // compute units of the device (persistent grid size)
  nwprt::DevBuf denoise_scratch;  // denoiser ping-pong images + normalisation
  std::map<int, std::unique_ptr<nwprt::TextureObj>> textures;
  std::map<int, std::unique_ptr<nwprt::Mesh>> meshes;
  nwprt::DevBuf env_tables;  // direct-light sampling tables of the last cast
  nwprt::DevBuf sh_response, sh_colors;  // SH-PRT render work buffers
  // SH transfer computation: the basis maps' sampling tables and the path
  // records, bases and per-(sample, map) radiance of one launch.
  nwprt::DevBuf sh_tables, sh_work;
  nwprt::DevBuf haar_work;  // nwprt_envmap_haar's tile scaling coefficients
};

namespace nwprt {

// Runs f, mapping nwprt::Error and other exceptions to a status code and
// the thread's last-error message (no exception crosses the C ABI).
template <class F>
nwprt_status guarded(F &&f) {
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

void bind(nwprt_ctx ctx);
hipStream_t pick(nwprt_ctx ctx, void *stream);
Scene &get_scene(nwprt_ctx ctx, int id);
EnvMap &get_env(nwprt_ctx ctx, int id);
void check_aligned(const void *p, size_t a, const char *what);
void check_frame(const nwprt_gbuffer *gb, int w, int h, bool need_direct);
void check_outputs(const nwprt_frame_out &out);
Elem elem_of(nwprt_precision p);
EncodeArgs encode_args(const Scene &sc, const nwprt_gbuffer *gb, int P, Elem elem,
                       void *features);
const void *scene_tiles(const Scene &sc, nwprt_precision prec);
// Whether a scene's per-pixel features take the planar layout for an MFMA
// element type (16 x 2 grids without aux inputs; the projection then moves
// into the selection tiles).
bool planar_features(const Scene &sc, Elem elem);
// Projection to fold into the selection tiles (null: none).
const float *scene_projection(const Scene &sc);
// Whether a relight of this scene at this precision encodes inside the
// fused kernel (MFMA forms, planar features), and the arguments that do it.
bool fuse_encode(const Scene &sc, nwprt_precision prec);
void set_fused_encode(RelightArgs &a, const Scene &sc, const nwprt_gbuffer *gb);
// Aux encodings of a scene with aux inputs: checks that the G-buffer holds
// them and points the encode at `aux` ([P][16] of the encode's element).
void set_aux_encode(EncodeArgs &e, const Scene &sc, const nwprt_gbuffer *gb, void *aux);
// Aux fields of a relight (features' aux inputs and the scene's aux tiles).
void set_aux_relight(RelightArgs &a, const Scene &sc, nwprt_precision prec, const void *aux);

// FP8 form of one relight: its work buffers, the per-relight preparation
// (selection tiles of both shapes, calibration sample -- encoded from the
// G-buffer for a fused relight, else gathered from a.features / a.aux --
// and activation maxima, on stream s), and its relight arguments.
struct Fp8Work {
  uint16_t *mk16 = nullptr;      // [K][4][64][8] BF16 16x16x32 tiles (calibration network)
  uint16_t *mk32 = nullptr;      // [K][2][2][64][8] BF16 32x32x16 tiles (FP8 first layer)
  uint16_t *calib = nullptr;     // calibration features, then (aux scenes) aux inputs
  float *amax = nullptr;         // 4 floats
};
size_t fp8_mk_bytes(int k);
size_t fp8_calib_bytes(const Scene &sc, int n_pixels);
void prepare_fp8(const Scene &sc, const int32_t *idx, int k, const RelightArgs &a,
                 const Fp8Work &w, hipStream_t s);
void set_fp8_relight(RelightArgs &a, const Scene &sc, const Fp8Work &w);

// Device work buffers of one relight, provided by the caller of
// relight_with: per-pixel features (unfused relights) and aux inputs
// (scenes with aux inputs), the MFMA selection tiles (FP8: the 16x16 tiles
// of the calibration network) or, for the plain FP32 form, the selected
// wavelet features; and the FP8 form's 32x32 tiles, calibration sample and
// maxima.  Sizes: mk K x 4096 B, wsel K x 128 B, mk32 fp8_mk_bytes(K),
// calib fp8_calib_bytes(scene, P), amax 16 B.
struct RelightWork {
  const void *features = nullptr;
  const void *aux = nullptr;
  uint16_t *mk = nullptr;
  float *wsel = nullptr;
  uint16_t *mk32 = nullptr, *calib = nullptr;
  float *amax = nullptr;
};
// One relight of P pixels on stream s with the given work buffers (no
// allocation): selection tiles, FP8 calibration where it applies, the
// relight kernel, and the map's last-use event.
void relight_with(nwprt_ctx ctx, const Scene &sc, EnvMap &env, const nwprt_gbuffer *gb, int P,
                  const nwprt_frame_out *out, nwprt_precision prec, nwprt_tonemap_source tm,
                  hipStream_t s, bool fused, const RelightWork &w, float *dump = nullptr,
                  int dump_begin = 0, int dump_count = 0, int variant = -1);

// The allocation-free entry point's per-device scratch array (noalloc.hip):
// resolves its address on the current device, once per device.
void scratch_init(int device);

}  // namespace nwprt
