// Allocation-free relight: nwprt_relight_noalloc.  No counterpart in the
// reference; it is for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  It runs the same kernels as nwprt_relight on the fused-encode
// forms (BF16, FP16, FP8 with 16 x 2 hash grids and no aux inputs), so its
// outputs equal nwprt_relight's bit for bit, but its work buffers -- the
// selection tiles, and for FP8 the 32x32 tiles, calibration sample and
// activation maxima -- live in one file-scope device array, `scratch`, of
// kScratchBytes, instead of the context's growable buffers.
//
// Each GPU holds its own copy of `scratch` until the process ends; its
// address on a device is resolved once, by nwprt_create with that device
// current (scratch_init), and kept per device.  Every call on a GPU shares
// that GPU's array: calls that overlap on one GPU (other streams, other
// contexts) race, and the array keeps the last call's contents.  Nothing in
// it must start at zero: the selection tiles are written whole, and the FP8
// calibration kernels clear the maxima before accumulating them.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine_state.hpp"
#include "kernels.hpp"

namespace nwprt {
namespace {

// K <= 1024: 4 MiB of 16x16 selection tiles; FP8 adds 4 MiB of 32x32 tiles,
// the calibration sample (at most 16384 pixels x 64 B = 1 MiB) and 16 B of
// maxima.
constexpr size_t kScratchBytes = 10u << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

void *scratch_on(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

}  // namespace

void scratch_init(int device) {
  require(device >= 0 && device < kMaxDevices, NWPRT_ERR_INVALID_ARGUMENT,
          "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  g_scratch_ptr[device] = p;
}

}  // namespace nwprt

using namespace nwprt;

extern "C" {

nwprt_status nwprt_relight_noalloc(nwprt_ctx ctx, int32_t scene_id, int32_t env_id,
                                   const nwprt_gbuffer *gbuffer, int32_t width, int32_t height,
                                   const nwprt_frame_out *out, nwprt_precision precision,
                                   nwprt_tonemap_source tonemap, void *stream) {
  return guarded([&] {
    bind(ctx);
    const Scene &sc = get_scene(ctx, scene_id);
    EnvMap &env = get_env(ctx, env_id);
    check_frame(gbuffer, width, height, true);
    require(precision == NWPRT_PREC_BF16 || precision == NWPRT_PREC_FP16 ||
                precision == NWPRT_PREC_FP8,
            NWPRT_ERR_UNSUPPORTED, "relight_noalloc: BF16, FP16 or FP8 precision only");
    require(fuse_encode(sc, precision), NWPRT_ERR_UNSUPPORTED,
            "relight_noalloc: the scene needs a 16 x 2 hash grid without aux inputs (the "
            "fused-encode form, which keeps no per-pixel buffers)");
    char *base = (char *)scratch_on(ctx->device);
    require(base != nullptr, NWPRT_ERR_STATE, "relight_noalloc: scratch not set up on this device");
    const int P = width * height;
    const size_t mk = round256((size_t)env.k * 4 * 64 * 16);
    size_t need = mk;
    RelightWork w;
    w.mk = (uint16_t *)base;
    if (precision == NWPRT_PREC_FP8) {
      const size_t mk32 = round256(fp8_mk_bytes(env.k)), cal = round256(fp8_calib_bytes(sc, P));
      w.mk32 = (uint16_t *)(base + mk);
      w.calib = (uint16_t *)(base + mk + mk32);
      w.amax = (float *)(base + mk + mk32 + cal);
      need = mk + mk32 + cal + 256;
    }
    require(need <= kScratchBytes, NWPRT_ERR_INVALID_ARGUMENT,
            "relight_noalloc: needs " + std::to_string(need) + " bytes of scratch, more than the " +
                std::to_string(kScratchBytes) + " it has");
    relight_with(ctx, sc, env, gbuffer, P, out, precision, tonemap, pick(ctx, stream), true, w);
  });
}

}  // extern "C"
