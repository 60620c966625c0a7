// valu_probe.hip -- sustained VALU issue rates of gfx950 for a few
// instruction mixes, the ceilings the pair kernels' VALU roofline is read
// against.
//
//   quest_amd/bin/valu_probe
//
// Every kernel runs 8 independent dependency chains per lane, 256-thread
// blocks, 8 blocks per CU (8 waves per SIMD), so issue, not latency, bounds
// it.  Printed per mix: wave-instructions per ns per CU and the lane-op rate.
// Mixes:
//   fma32_sgpr  v_fma_f32 with a scalar operand (single issue only)
//   fmac32_vgpr v_fmac_f32, all operands in VGPRs (may dual-issue)
//   mul32_vgpr  v_mul_f32, all operands in VGPRs
//   fma64       v_fma_f64
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int kThreads = 256;
constexpr int kChains = 8;

template <int KIND>
__global__ __launch_bounds__(kThreads) void probe(float *out32, double *out64, int iters, float a)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const float y = static_cast<float>(threadIdx.x) * 1e-7f;          // per-lane (VGPR) operands
  const float z = 0.999f - static_cast<float>(threadIdx.x) * 1e-9f;
  if constexpr(KIND == 3) {
    double x[kChains];
    for(int c = 0; c < kChains; ++c) x[c] = t + c;
#pragma unroll 4
    for(int i = 0; i < iters; ++i)
#pragma unroll
      for(int c = 0; c < kChains; ++c) x[c] = __builtin_fma(x[c], static_cast<double>(z), static_cast<double>(y));
    double s = 0;
    for(int c = 0; c < kChains; ++c) s += x[c];
    out64[t] = s;
  } else {
    float x[kChains];
    for(int c = 0; c < kChains; ++c) x[c] = t + c;
#pragma unroll 4
    for(int i = 0; i < iters; ++i)
#pragma unroll
      for(int c = 0; c < kChains; ++c) {
        if constexpr(KIND == 0) x[c] = __builtin_fmaf(x[c], a, 0.001f);  // scalar operand
        if constexpr(KIND == 1) x[c] = __builtin_fmaf(y, z, x[c]);       // v_fmac_f32, VGPRs
        if constexpr(KIND == 2) x[c] = x[c] * z;                          // v_mul_f32, VGPRs
      }
    float s = 0;
    for(int c = 0; c < kChains; ++c) s += x[c];
    out32[t] = s;
  }
}

#define CHECK(call)                                                                   \
  do {                                                                                \
    const hipError_t e_ = (call);                                                     \
    if(e_ != hipSuccess) {                                                            \
      std::fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));                 \
      return 1;                                                                       \
    }                                                                                 \
  } while(0)

}  // namespace

int main()
{
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  const int cus = prop.multiProcessorCount;
  const int blocks = cus * 8, n = blocks * kThreads, iters = 20000;
  float *o32 = nullptr;
  double *o64 = nullptr;
  CHECK(hipMalloc(&o32, sizeof(float) * n));
  CHECK(hipMalloc(&o64, sizeof(double) * n));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  std::printf("%s, %d CUs, peak clock %.0f MHz\n", prop.gcnArchName, cus, prop.clockRate / 1e3);
  const char *names[] = {"fma32_sgpr", "fmac32_vgpr", "mul32_vgpr", "fma64"};
  for(int kind = 0; kind < 4; ++kind) {
    float best = 1e30f;
    for(int rep = 0; rep < 3; ++rep) {
      CHECK(hipEventRecord(e0));
      switch(kind) {
        case 0: hipLaunchKernelGGL(probe<0>, blocks, kThreads, 0, 0, o32, o64, iters, 0.999f); break;
        case 1: hipLaunchKernelGGL(probe<1>, blocks, kThreads, 0, 0, o32, o64, iters, 0.999f); break;
        case 2: hipLaunchKernelGGL(probe<2>, blocks, kThreads, 0, 0, o32, o64, iters, 0.999f); break;
        default: hipLaunchKernelGGL(probe<3>, blocks, kThreads, 0, 0, o32, o64, iters, 0.999f); break;
      }
      CHECK(hipGetLastError());
      CHECK(hipEventRecord(e1));
      CHECK(hipEventSynchronize(e1));
      float ms = 0;
      CHECK(hipEventElapsedTime(&ms, e0, e1));
      if(ms < best) best = ms;
    }
    const double wave_instr = static_cast<double>(kChains) * iters * (n / 64.0);
    const double per_ns_cu = wave_instr / (best * 1e6) / cus;
    std::printf("%-12s %8.3f ms  %.3f wave-instr/ns/CU  %.2f T lane-ops/s\n", names[kind], best, per_ns_cu,
                wave_instr * 64 / (best * 1e-3) / 1e12);
  }
  CHECK(hipFree(o32));
  CHECK(hipFree(o64));
  return 0;
}
