// valu_peak [iters]: the FP32 VALU issue rate a CU of this device sustains, the
// ceiling against which the integrator and filter kernels' VALU issue rates are
// read.
//
// k_fma: every lane runs 8 independent chains of fused multiply-adds (no
// dependency stalls at 8 wavefronts per SIMD), so the loop is nothing but
// v_fma_f32 (v_pk_fma_f32 is disabled at build time, so each is one wavefront
// instruction).  k_mix: the same with every second chain an exp2 (a quarter-rate
// transcendental).  Each kernel runs twice (the first launch includes clock
// ramp-up); the program prints wavefront instructions of the loop per CU per
// nanosecond.  With rocprofv3 --pmc SQ_INSTS_VALU GRBM_GUI_ACTIVE the counters
// give the same figure per cycle: SQ_INSTS_VALU / (CUs x GRBM_GUI_ACTIVE / XCDs).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kBlocks = 256 * 8;  // 8 blocks of 256 lanes per CU on a 256-CU device
constexpr int kThreads = 256;
constexpr int kChains = 8;

__global__ void __launch_bounds__(kThreads) k_fma(float* __restrict__ out, int iters) {
  float a[kChains];
#pragma unroll
  for (int k = 0; k < kChains; ++k) a[k] = (float)(threadIdx.x + k) * 1e-3f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = __builtin_fmaf(a[k], 0.9999f, 1e-4f);
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < kChains; ++k) s += a[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(kThreads) k_mix(float* __restrict__ out, int iters) {
  float a[kChains];
#pragma unroll
  for (int k = 0; k < kChains; ++k) a[k] = (float)(threadIdx.x + k) * 1e-3f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kChains; k += 2) {
      a[k] = __builtin_fmaf(a[k], 0.9999f, 1e-4f);
      a[k + 1] = __builtin_amdgcn_exp2f(a[k + 1] * -0.5f);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < kChains; ++k) s += a[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

#define CHECK(x)                                                                 \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));       \
      std::exit(1);                                                              \
    }                                                                            \
  } while (0)

}  // namespace

int main(int argc, char** argv) {
  const int iters = argc > 1 ? std::atoi(argv[1]) : 20000;
  if (iters <= 0 || iters > 1000000) {
    std::fprintf(stderr, "iters must be in [1, 1000000]\n");
    return 2;
  }
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  const int cus = prop.multiProcessorCount;
  float* d = nullptr;
  CHECK(hipMalloc(&d, (size_t)kBlocks * kThreads * sizeof(float)));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  // wavefront instructions of the loops: fma chains in k_fma; fma + exp2 + the
  // exp2's argument multiply in k_mix
  const double waves = (double)kBlocks * (kThreads / 64);
  const double inst_fma = waves * iters * kChains;
  const double inst_mix = waves * iters * (kChains / 2) * 3;
  std::printf("device %s, %d CUs, %d blocks of %d lanes, %d iterations\n", prop.name, cus, kBlocks, kThreads, iters);
  for (int rep = 0; rep < 2; ++rep) {
    float ms = 0.f;
    CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL(k_fma, dim3(kBlocks), dim3(kThreads), 0, 0, d, iters);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("k_fma run %d: %.3f ms, %.4e wavefront instructions, %.3f per CU per ns\n", rep, ms, inst_fma,
                inst_fma / cus / (ms * 1e6));
    CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL(k_mix, dim3(kBlocks), dim3(kThreads), 0, 0, d, iters);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("k_mix run %d: %.3f ms, %.4e wavefront instructions, %.3f per CU per ns\n", rep, ms, inst_mix,
                inst_mix / cus / (ms * 1e6));
  }
  float h = 0.f;
  CHECK(hipMemcpy(&h, d, sizeof(float), hipMemcpyDeviceToHost));
  std::printf("checksum %g\n", (double)h);
  CHECK(hipFree(d));
  return 0;
}
