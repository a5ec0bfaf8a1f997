// nwprt_render_engine.hpp -- RenderEngine-shaped C++ facade over the C ABI.
//
// Header-only; link libnwprt.so and the HIP runtime.  It keeps the method
// names and argument meaning of the reference's TCNN interface
// (Renderer/RenderEngine.h:78-85) so a renderer written against it can swap
// implementations: two frame-buffer sets as fullFrameCUDA holds them
// (Renderer/frames.h:96-183), initTCNN(modelPath), buildLights from
// env-map folders, TCNNPass_p1 / TCNNPass_p2 / TCNNSync /
// computeFinalPixelColors.  Errors throw std::runtime_error (the reference's
// CUDA_CHECK convention, Renderer/optix7.h:26-37); a bad env-map id throws
// std::overflow_error (Renderer/RenderEngine.cpp:348-349, here with the
// bound checked as >=).
//
// The OptiX pass that fills the G-buffer is not part of this build: the
// caller (or writeSyntheticGBuffer) writes fbFirstHitPos & co.  The
// denoiser step (RenderEngine::denoiseOrMoveData, RenderEngine.cpp:390-403)
// runs nwprt_denoise -- a guided a-trous filter in place of the OptiX
// learned denoiser -- with the frame's world normals and diffuse albedo
// (fbFirstHitKd) as guides, or copies when the denoiser is off.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "nwprt.h"

namespace nwprt {

enum class finalColorBufferContent { full, direct, indirect };

// Renderer/RenderEngine.h:29-42 (the fields the relight path reads) plus
// the MLP precision of this build.
struct renderConfig {
  finalColorBufferContent renderingMode = finalColorBufferContent::full;
  bool waveletMode = true;
  bool finalDenoiser = false;
  bool indirectDenoiser = false;
  nwprt_precision precision = NWPRT_PREC_BF16;
};

// Device buffers of one frame that the relight path reads or writes.
struct fullFrameBuffers {
  float *fbFirstHitPos = nullptr, *fbworldNormal = nullptr, *fbFirstHitReflecDir = nullptr,
        *fbRoughness = nullptr, *fbFirstHitKd = nullptr, *fbFirstHitKs = nullptr,
        *fbDirectDenoised = nullptr;
  uint8_t *fbEnvironmentMapFlag = nullptr;
  float *fbPredictedIndirect = nullptr, *fbPredictedIndirectDenoised = nullptr,
        *fbPredictedIndirectWithDirect = nullptr, *fbPredictedIndirectWithDirectDenoised = nullptr;
  uint32_t *fbFinalColor = nullptr;
};

class RenderEngine {
 public:
  explicit RenderEngine(int device = 0) : device_(device) { check(nwprt_create(device, &ctx_)); }
  ~RenderEngine() {
    freeFrames();
    if (ctx_) nwprt_destroy(ctx_);
  }
  RenderEngine(const RenderEngine &) = delete;
  RenderEngine &operator=(const RenderEngine &) = delete;

  // RenderEngine::resize: (re)allocate both frame-buffer sets.
  void resize(int width, int height) {
    if (width <= 0 || height <= 0) throw std::runtime_error("resize: bad frame size");
    freeFrames();
    width_ = width;
    height_ = height;
    const size_t P = (size_t)width * height;
    for (auto &f : frames_) {
      float **f4[] = {&f.fbFirstHitPos, &f.fbworldNormal, &f.fbFirstHitReflecDir,
                      &f.fbFirstHitKd, &f.fbFirstHitKs, &f.fbDirectDenoised,
                      &f.fbPredictedIndirect, &f.fbPredictedIndirectDenoised,
                      &f.fbPredictedIndirectWithDirect,
                      &f.fbPredictedIndirectWithDirectDenoised};
      for (float **b : f4) *b = (float *)alloc(P * 16);
      f.fbRoughness = (float *)alloc(P * 4);
      f.fbEnvironmentMapFlag = (uint8_t *)alloc(P);
      f.fbFinalColor = (uint32_t *)alloc(P * 4);
    }
  }
  int width() const { return width_; }
  int height() const { return height_; }
  fullFrameBuffers &frame(int frameID) {
    if (frameID < 0 || frameID > 1) throw std::out_of_range("frameID must be 0 or 1");
    return frames_[frameID];
  }

  // Fill a frame's G-buffer with the seeded synthetic generator (stands in
  // for the OptiX pass, Renderer/device_functions/camera.cu:247-270).
  void writeSyntheticGBuffer(int frameID, uint64_t seed, float backgroundFraction = 0.f) {
    const size_t P = (size_t)width_ * height_;
    std::vector<float> pos(4 * P), nrm(4 * P), refl(4 * P), kd(4 * P), ks(4 * P), dir(4 * P),
        rough(P);
    std::vector<uint8_t> flag(P);
    nwprt_gbuffer g{pos.data(), nrm.data(), refl.data(), rough.data(),
                    kd.data(), ks.data(), dir.data(), flag.data()};
    check(nwprt_synth_gbuffer(seed, width_, height_, backgroundFraction, &g));
    fullFrameBuffers &f = frame(frameID);
    upload(f.fbFirstHitPos, pos.data(), P * 16);
    upload(f.fbworldNormal, nrm.data(), P * 16);
    upload(f.fbFirstHitReflecDir, refl.data(), P * 16);
    upload(f.fbFirstHitKd, kd.data(), P * 16);
    upload(f.fbFirstHitKs, ks.data(), P * 16);
    upload(f.fbDirectDenoised, dir.data(), P * 16);
    upload(f.fbRoughness, rough.data(), P * 4);
    upload(f.fbEnvironmentMapFlag, flag.data(), P);
  }

  // Scene geometry for render(): a mesh from nwprt_mesh_create, and the
  // bounds that normalise first-hit positions (scene->bounds,
  // Renderer/RenderEngine.cpp:104-105); null bounds: the mesh's own.
  void setMesh(int32_t meshId, const float *bounds = nullptr) {
    mesh_ = meshId;
    hasBounds_ = bounds != nullptr;
    if (bounds)
      for (int i = 0; i < 6; ++i) bounds_[i] = bounds[i];
  }
  // RenderEngine::setCamera (Renderer/RenderEngine.cpp:246-267).
  void setCamera(const nwprt_camera &camera) { camera_ = camera; }

  // RenderEngine::render (Renderer/RenderEngine.cpp:328-353), G-buffer part
  // in wavelet mode (maximumBounce 1): first-hit encodings of frameID from
  // the mesh and camera.  viewEnvFaces (device [6][n][n][3], optional) gives
  // the kd of pixels that see the environment.  With lightEnvFaces the
  // direct-light buffer gets `directSamples` estimates per pixel
  // (samplesPerPixel x directSamplesPerLight of renderConfig); without it,
  // that buffer is not written.
  void render(int frameID, hipStream_t stream = nullptr, const float *viewEnvFaces = nullptr,
              int viewEnvN = 0, const float *lightEnvFaces = nullptr, int lightEnvN = 0,
              int directSamples = 1, uint32_t seed = 0) {
    nwprt_gbuffer g = gbuffer(frameID);
    nwprt_cast_params p{1,          hasBounds_ ? bounds_ : nullptr, viewEnvFaces, viewEnvN,
                        lightEnvFaces, lightEnvN,                   directSamples, seed};
    check(nwprt_gbuffer_cast(ctx_, mesh_, &camera_, width_, height_, &p, &g, stream));
  }

  // Zero frameID's direct-light buffer (for G-buffers from render() without
  // a lighting map, which leaves that buffer as it was).
  void clearDirectLight(int frameID) {
    hipCheck(hipMemset(frame(frameID).fbDirectDenoised, 0, (size_t)width_ * height_ * 16));
  }

  // RenderEngine::initTCNN (Renderer/tinycudann/tcnn.cu:49-81).
  void initTCNN(const std::string &modelPath) {
    int32_t id = -1;
    check(nwprt_scene_load_folder(ctx_, modelPath.c_str(), &id));
    scene_ = id;
  }
  // Use a scene created through the C ABI instead of a weight folder.
  void setScene(int32_t sceneId) { scene_ = sceneId; }
  nwprt_ctx context() const { return ctx_; }

  // RenderEngine::buildLights, wavelet part (Renderer/RenderEngine.cpp:
  // 113-175): one Environment_Map(folder, name) per entry; maps without a
  // stored selection get a GPU transform with `k` wavelets.
  void buildLights(const std::vector<std::pair<std::string, std::string>> &envmaps, int k) {
    destroyLights();
    for (const auto &e : envmaps) {
      int32_t id = -1;
      check(nwprt_envmap_load_folder(ctx_, e.first.c_str(), e.second.c_str(), k, &id));
      envs_.push_back(id);
    }
  }
  int addEnvironmentMap(int32_t envId) {
    envs_.push_back(envId);
    return (int)envs_.size() - 1;
  }
  void destroyLights() {
    for (int32_t id : envs_) nwprt_envmap_destroy(ctx_, id);
    envs_.clear();
  }
  size_t environmentMapCount() const { return envs_.size(); }

  // RenderEngine::TCNNAvaliable (tcnn.cu:160-162).
  bool TCNNAvaliable() const { return scene_ >= 0 && !envs_.empty(); }

  // RenderEngine::TCNNPass_p1 (tcnn.cu:83-108): per-pixel, env-independent
  // part of the network for frameID.
  void TCNNPass_p1(int frameID, hipStream_t stream = nullptr, renderConfig config = {}) {
    nwprt_gbuffer g = gbuffer(frameID);
    check(nwprt_pass_p1(ctx_, scene_, frameID, &g, width_, height_, config.precision, stream));
  }

  // RenderEngine::TCNNPass_p2 (tcnn.cu:111-152): per-pair network, clamped
  // K-sum, then the denoise-or-copy into the "Denoised" buffers
  // (tcnn.cu:144-148: finalDenoiser for the with-direct image,
  // indirectDenoiser for the indirect one, blend factor 0).
  void TCNNPass_p2(int envMapID, int frameID, renderConfig config, hipStream_t stream = nullptr) {
    if (envMapID < 0 || (size_t)envMapID >= envs_.size())
      throw std::overflow_error("Envmap ID for rendering is larger than cached envmap");
    nwprt_gbuffer g = gbuffer(frameID);
    fullFrameBuffers &f = frame(frameID);
    nwprt_frame_out out{f.fbPredictedIndirect, f.fbPredictedIndirectWithDirect, nullptr};
    check(nwprt_pass_p2(ctx_, scene_, envs_[(size_t)envMapID], frameID, &g, width_, height_, &out,
                        config.precision, NWPRT_TONEMAP_NONE, stream));
    denoiseOrMoveData(frameID, config.finalDenoiser, f.fbPredictedIndirectWithDirect,
                      f.fbPredictedIndirectWithDirectDenoised, 0.f, stream);
    denoiseOrMoveData(frameID, config.indirectDenoiser, f.fbPredictedIndirect,
                      f.fbPredictedIndirectDenoised, 0.f, stream);
    denoised_ = config.finalDenoiser || config.indirectDenoiser;
  }

  // RenderEngine::denoiseOrMoveData (RenderEngine.cpp:390-403): denoise
  // `input` into `output` guided by frameID's normals and albedo, or copy.
  void denoiseOrMoveData(int frameID, bool denoiserOn, const float *input, float *output,
                         float blendFactor = 0.f, hipStream_t stream = nullptr) {
    fullFrameBuffers &f = frame(frameID);
    if (denoiserOn) {
      check(nwprt_denoise(ctx_, input, f.fbworldNormal, f.fbFirstHitKd, output, width_, height_,
                          blendFactor, &denoiseParams_, stream));
    } else {
      hipCheck(hipMemcpyAsync(output, input, (size_t)width_ * height_ * 16,
                              hipMemcpyDeviceToDevice, stream ? stream : ownStream()));
    }
  }
  // RenderEngine::denoise (RenderEngine.cpp:405-408) on caller images.
  void denoise(const float *color, const float *normal, const float *albedo, float *output,
               int width, int height, hipStream_t stream = nullptr) {
    check(nwprt_denoise(ctx_, color, normal, albedo, output, width, height, 0.f,
                        &denoiseParams_, stream));
  }
  void setDenoiseParams(const nwprt_denoise_params &p) { denoiseParams_ = p; }

  // RenderEngine::TCNNSync (tcnn.cu:154-158).
  void TCNNSync() {
    check(nwprt_sync(ctx_));
    hipCheck(hipDeviceSynchronize());
  }

  // RenderEngine::computeFinalPixelColors, wavelet mode (toneMap.cu:73-99).
  void computeFinalPixelColors(int frameID = 1, hipStream_t stream = nullptr,
                               renderConfig config = {}) {
    fullFrameBuffers &f = frame(frameID);
    const float *src = config.renderingMode == finalColorBufferContent::indirect
                           ? f.fbPredictedIndirectDenoised
                           : f.fbPredictedIndirectWithDirectDenoised;
    if (config.renderingMode == finalColorBufferContent::direct) src = f.fbDirectDenoised;
    check(nwprt_tonemap(ctx_, src, f.fbFinalColor, width_, height_, stream));
  }

  // RenderEngine::downloadPixels (Renderer/RenderEngine.cpp:286-291).
  void downloadPixels(uint32_t *hostPixels, int frameID) {
    TCNNSync();
    hipCheck(hipMemcpy(hostPixels, frame(frameID).fbFinalColor,
                       (size_t)width_ * height_ * 4, hipMemcpyDeviceToHost));
  }
  // The two wavelet outputs of RenderEngine::downloadframe (RenderEngine.cpp:317-318).
  void downloadPredictions(int frameID, std::vector<float> &indirect,
                           std::vector<float> &withDirect) {
    TCNNSync();
    const size_t n = (size_t)width_ * height_ * 4;
    indirect.resize(n);
    withDirect.resize(n);
    hipCheck(hipMemcpy(indirect.data(), frame(frameID).fbPredictedIndirectDenoised, n * 4,
                       hipMemcpyDeviceToHost));
    hipCheck(hipMemcpy(withDirect.data(), frame(frameID).fbPredictedIndirectWithDirectDenoised,
                       n * 4, hipMemcpyDeviceToHost));
  }
  // Whether the last TCNNPass_p2 ran the denoiser on either output.
  bool denoiserApplied() const { return denoised_; }

 private:
  static void check(nwprt_status s) {
    if (s != NWPRT_OK)
      throw std::runtime_error(std::string("nwprt: ") + nwprt_status_string(s) + ": " +
                               nwprt_last_error());
  }
  static void hipCheck(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e));
  }
  void *alloc(size_t bytes) {
    void *p = nullptr;
    hipCheck(hipSetDevice(device_));
    hipCheck(hipMalloc(&p, bytes));
    allocations_.push_back(p);
    return p;
  }
  static void upload(void *dst, const void *src, size_t bytes) {
    hipCheck(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  }
  void freeFrames() {
    if (!allocations_.empty()) (void)hipDeviceSynchronize();
    for (void *p : allocations_) (void)hipFree(p);
    allocations_.clear();
    frames_[0] = frames_[1] = fullFrameBuffers{};
  }
  nwprt_gbuffer gbuffer(int frameID) {
    fullFrameBuffers &f = frame(frameID);
    return nwprt_gbuffer{f.fbFirstHitPos, f.fbworldNormal, f.fbFirstHitReflecDir, f.fbRoughness,
                         f.fbFirstHitKd,  f.fbFirstHitKs,  f.fbDirectDenoised,    f.fbEnvironmentMapFlag};
  }
  hipStream_t ownStream() {
    void *s = nullptr;
    check(nwprt_get_stream(ctx_, &s));
    return (hipStream_t)s;
  }

  int device_;
  nwprt_ctx ctx_ = nullptr;
  int32_t scene_ = -1;
  int32_t mesh_ = -1;
  nwprt_camera camera_{{0, 0, 0}, {0, 0, -1}, {0, 1, 0}};
  float bounds_[6] = {0, 0, 0, 0, 0, 0};
  bool hasBounds_ = false;
  std::vector<int32_t> envs_;
  int width_ = 0, height_ = 0;
  fullFrameBuffers frames_[2];
  std::vector<void *> allocations_;
  nwprt_denoise_params denoiseParams_ = defaultDenoiseParams();
  bool denoised_ = false;
  static nwprt_denoise_params defaultDenoiseParams() {
    nwprt_denoise_params p;
    nwprt_denoise_default_params(&p);
    return p;
  }
};

}  // namespace nwprt
