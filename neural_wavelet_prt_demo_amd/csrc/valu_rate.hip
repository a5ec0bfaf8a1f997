// Issue cost of the vector-ALU instructions the relight kernels' ReLU +
// convert and output steps are made of: the read-out behind the FP8 form's
// limit.  One kernel per instruction: every lane runs eight independent
// chains of that one instruction (no memory traffic besides one store per
// lane at the end), eight waves per SIMD, so the SIMD issues nothing else.
// Prints one JSON line per instruction: the median of five launches, the
// time per wave64 instruction per SIMD in cycles of the GPU's maximum clock,
// and the ratio to v_cvt_pk_f16_f32 (a full-rate conversion), which does
// not depend on the clock the GPU actually ran at.  Built next to libnwprt
// as `valu_rate`; reads and writes nothing but its own device buffer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <vector>

namespace {

constexpr int kThreads = 256, kIters = 2048, kChains = 8, kWavesPerSimd = 8;

// Eight chains: r[i] <- op(r[i], r[i + 1]); the compiler keeps the asm as
// written (volatile) and the values feed the final store.
#define NWPRT_CHAINS(ASM)                                                          \
  uint32_t r[kChains];                                                             \
  for (int i = 0; i < kChains; ++i)                                                \
    r[i] = (threadIdx.x * 2654435761u) ^ (i * 40503u) ^ 0x3c003c00u;               \
  const float sc = 0.00390625f;                                                    \
  for (int it = 0; it < iters; ++it) {                                             \
    _Pragma("unroll") for (int i = 0; i < kChains; ++i) { ASM; }                   \
  }                                                                                \
  uint32_t x = 0;                                                                  \
  for (int i = 0; i < kChains; ++i) x ^= r[i];                                     \
  out[blockIdx.x * kThreads + threadIdx.x] = x;

#define NWPRT_BINARY(NAME, TEXT)                                                   \
  __global__ __launch_bounds__(kThreads) void NAME(uint32_t *out, int iters) {     \
    NWPRT_CHAINS(asm volatile(TEXT : "=v"(r[i]) : "v"(r[i]), "v"(r[(i + 1) & 7]))) \
  }

NWPRT_BINARY(k_cvt_pk_bf16, "v_cvt_pk_bf16_f32 %0, %1, %2")
NWPRT_BINARY(k_pk_max_i16, "v_pk_max_i16 %0, %1, %2")
NWPRT_BINARY(k_cvt_pk_f16_clamp, "v_cvt_pk_f16_f32 %0, %1, %2 clamp")
NWPRT_BINARY(k_cvt_pk_fp8, "v_cvt_pk_fp8_f32 %0, %1, %2")
NWPRT_BINARY(k_add_f32, "v_add_f32 %0, %1, %2")

// The scaled FP8 conversion also reads its destination (the word it keeps).
__global__ __launch_bounds__(kThreads) void k_cvt_scalef32_fp8_f16(uint32_t *out, int iters) {
  NWPRT_CHAINS(asm volatile("v_cvt_scalef32_pk_fp8_f16 %0, %1, %2"
                            : "+v"(r[i])
                            : "v"(r[(i + 1) & 7]), "v"(sc)))
}
__global__ __launch_bounds__(kThreads) void k_med3_f32(uint32_t *out, int iters) {
  NWPRT_CHAINS(asm volatile("v_med3_f32 %0, %1, %2, %3"
                            : "=v"(r[i])
                            : "v"(r[i]), "v"(r[(i + 1) & 7]), "v"(sc)))
}
__global__ __launch_bounds__(kThreads) void k_exp_f32(uint32_t *out, int iters) {
  NWPRT_CHAINS(asm volatile("v_exp_f32 %0, %1" : "=v"(r[i]) : "v"(r[i])))
}
__global__ __launch_bounds__(kThreads) void k_log_f32(uint32_t *out, int iters) {
  NWPRT_CHAINS(asm volatile("v_log_f32 %0, %1" : "=v"(r[i]) : "v"(r[i])))
}

bool ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
  return false;
}

}  // namespace

int main() {
  hipDeviceProp_t prop;
  if (!ok(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties")) return 1;
  const int n_simd = prop.multiProcessorCount * 4;
  const int blocks = n_simd * kWavesPerSimd * 64 / kThreads;
  uint32_t *out = nullptr;
  if (!ok(hipMalloc(&out, (size_t)blocks * kThreads * sizeof(uint32_t)), "hipMalloc")) return 1;
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "hipEventCreate") || !ok(hipEventCreate(&e1), "hipEventCreate"))
    return 1;
  struct Op {
    const char *name;
    void (*kernel)(uint32_t *, int);
  };
  const Op ops[] = {{"v_cvt_pk_f16_f32 (clamp)", k_cvt_pk_f16_clamp},
                    {"v_cvt_pk_bf16_f32", k_cvt_pk_bf16},
                    {"v_pk_max_i16", k_pk_max_i16},
                    {"v_med3_f32", k_med3_f32},
                    {"v_add_f32", k_add_f32},
                    {"v_cvt_scalef32_pk_fp8_f16", k_cvt_scalef32_fp8_f16},
                    {"v_cvt_pk_fp8_f32", k_cvt_pk_fp8},
                    {"v_exp_f32", k_exp_f32},
                    {"v_log_f32", k_log_f32}};
  const double max_clock_hz = prop.clockRate * 1e3;
  const double inst_per_simd = (double)kWavesPerSimd * kIters * kChains;
  double ref_ms = 0.0;
  for (const Op &op : ops) {
    std::vector<float> t;
    for (int rep = 0; rep < 5; ++rep) {
      if (!ok(hipEventRecord(e0, nullptr), "hipEventRecord")) return 1;
      hipLaunchKernelGGL(op.kernel, dim3(blocks), dim3(kThreads), 0, nullptr, out, kIters);
      if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(e1, nullptr), "hipEventRecord") ||
          !ok(hipEventSynchronize(e1), "hipEventSynchronize"))
        return 1;
      float ms = 0.f;
      if (!ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) return 1;
      t.push_back(ms);
    }
    std::sort(t.begin(), t.end());
    const double ms = t[t.size() / 2];
    if (ref_ms == 0.0) ref_ms = ms;
    std::printf("{\"instruction\": \"%s\", \"ms\": %.4f, \"cycles_per_wave64_at_max_clock\": %.2f, "
                "\"relative_to_cvt_pk_f16\": %.2f, \"max_clock_ghz\": %.2f}\n",
                op.name, ms, ms * 1e-3 * max_clock_hz / inst_per_simd, ms / ref_ms,
                max_clock_hz / 1e9);
  }
  return hipFree(out) == hipSuccess ? 0 : 1;
}
