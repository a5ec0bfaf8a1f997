// Sustained v_mfma_f32_16x16x32_bf16 rate of this GPU: the ceiling against
// which the fused relight kernel's MFMA rate is read.
//
// Each wave keeps eight independent accumulator chains busy for ~25 ms per
// launch (long enough for the clock to settle under load), optionally with
// NV independent FP32 FMAs scheduled after each MFMA.  Two operand kinds:
// - constant: one A and one B vector for every MFMA (little switching);
// - random: eight A and eight B vectors of random finite bf16 values,
//   cycled so consecutive MFMAs read different data, as a network's
//   weights and activations do.  The clock under this load is the one a
//   real MFMA-bound kernel gets.
// Prints one JSON line per configuration.  Built next to libnwprt as
// `mfma_ceiling`; reads and writes nothing but its own device buffer.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;
constexpr int kChains = 8;

template <bool RANDOM, int NV>
__global__ __launch_bounds__(kBlock, 2) void mfma_loop(float *out, int iters) {
  unsigned x = (blockIdx.x * kBlock + threadIdx.x) * 2654435761u + 12345u;
  auto next = [&]() {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
  };
  bf16x8 a[kChains], b[kChains];
#pragma unroll
  for (int i = 0; i < kChains; ++i) {
    u32x4 ua, ub;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      // Random sign and mantissa, exponent in [2^-15, 2): finite, no denormals.
      ua[j] = RANDOM ? (next() & 0xBFFFBFFFu) | 0x38003800u : 0x3f803f80u;
      ub[j] = RANDOM ? (next() & 0xBFFFBFFFu) | 0x38003800u : 0x3f803f80u;
    }
    a[i] = __builtin_bit_cast(bf16x8, ua);
    b[i] = __builtin_bit_cast(bf16x8, ub);
  }
  f32x4 c[kChains];
  float v[kChains];
#pragma unroll
  for (int i = 0; i < kChains; ++i) {
    c[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    v[i] = (float)(threadIdx.x + i);
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int i = 0; i < kChains; ++i) {
        const bf16x8 ai = a[(i + 5 * h) % kChains], bi = b[(i + 3 + 5 * h) % kChains];
        c[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ai, bi, c[i], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < NV; ++j)
          v[(i + j) % kChains] = __builtin_fmaf(v[(i + j) % kChains], 1.0001f, 0.5f);
        if (NV > 0) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
          __builtin_amdgcn_sched_group_barrier(0x002, NV, 0);  // then NV VALU
        }
      }
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < kChains; ++i) s += c[i][0] + c[i][1] + c[i][2] + c[i][3] + v[i];
  out[blockIdx.x * kBlock + threadIdx.x] = s;
}

int check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return 1;
  }
  return 0;
}

template <bool RANDOM, int NV>
int run(int blocks, int iters, float *d) {
  hipEvent_t e0, e1;
  if (check(hipEventCreate(&e0), "hipEventCreate") || check(hipEventCreate(&e1), "hipEventCreate"))
    return 1;
  const int reps = 8;
  hipLaunchKernelGGL((mfma_loop<RANDOM, NV>), dim3(blocks), dim3(kBlock), 0, 0, d, iters);
  if (check(hipGetLastError(), "launch")) return 1;
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL((mfma_loop<RANDOM, NV>), dim3(blocks), dim3(kBlock), 0, 0, d, iters);
  (void)hipEventRecord(e1, 0);
  if (check(hipEventSynchronize(e1), "hipEventSynchronize")) return 1;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  const double flop = (double)blocks * (kBlock / 64) * iters * 2 * kChains * 16384.0 * reps;
  std::printf("{\"mfma\": \"v_mfma_f32_16x16x32_bf16\", \"operands\": \"%s\", "
              "\"valu_per_mfma\": %d, \"tflops\": %.1f, \"ms_per_launch\": %.3f}\n",
              RANDOM ? "random" : "constant", NV, flop / (ms * 1e-3) / 1e12, ms / reps);
  std::fflush(stdout);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

}  // namespace

int main() {
  int cus = 0;
  if (check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0),
            "hipDeviceGetAttribute"))
    return 1;
  // Eight 4-wave workgroups per CU = eight waves per SIMD, eight chains each.
  const int blocks = cus * 8, iters = 25000;
  float *d = nullptr;
  if (check(hipMalloc(&d, (size_t)blocks * kBlock * sizeof(float)), "hipMalloc")) return 1;
  int rc = run<false, 0>(blocks, iters, d);
  rc = rc ? rc : run<true, 0>(blocks, iters, d);
  rc = rc ? rc : run<true, 2>(blocks, iters, d);
  rc = rc ? rc : run<true, 3>(blocks, iters, d);
  (void)hipFree(d);
  return rc;
}
