// mfma_ceiling.hip -- the MFMA rate this part sustains with no memory traffic:
// every wave issues independent v_mfma_f64_16x16x4f64 (or f32) on register
// operands, 8 accumulators per wave, 1/2/4 workgroups of 4 waves per CU.  The
// result is the practical ceiling for the FP64 and FP32 MFMA kernels of the
// solver (dense filter step, Gram products, tall-skinny products).
//
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops \
//       rayleighchebyshevgpgpu_amd/probes/mfma_ceiling.hip -o /tmp/mfma_ceiling && /tmp/mfma_ceiling
#include <hip/hip_runtime.h>

#include <cstdio>

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int kAcc = 8;
constexpr int kThreads = 256;

// out has one entry per thread of the grid; nothing else is read or written
__global__ __launch_bounds__(kThreads) void mfma_f64_loop(double* out, int iters, double a0) {
  d4v acc[kAcc];
  for (int i = 0; i < kAcc; ++i) acc[i] = d4v{0.0, 0.0, 0.0, 0.0};
  const double a = a0 + threadIdx.x * 1e-9, b = 0.5 * a0;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < kAcc; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
  }
  double s = 0.0;
  for (int i = 0; i < kAcc; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * kThreads + threadIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void mfma_f32_loop(float* out, int iters, float a0) {
  f4v acc[kAcc];
  for (int i = 0; i < kAcc; ++i) acc[i] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
  const float a = a0 + threadIdx.x * 1e-6f, b = 0.5f * a0;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < kAcc; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i], 0, 0, 0);
  }
  float s = 0.0f;
  for (int i = 0; i < kAcc; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * kThreads + threadIdx.x] = s;
}

#define CHECK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                               \
    }                                                                         \
  } while (0)

int main() {
  int dev = 0, ncu = 0;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int max_wg_per_cu = 4;
  void* out = nullptr;
  CHECK(hipMalloc(&out, sizeof(double) * (size_t)max_wg_per_cu * ncu * kThreads));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  const int iters = 20000;
  std::printf("%d CUs, %d MFMAs of 16x16x4 per wave, 4 waves per workgroup\n", ncu, iters * kAcc);
  for (int wg : {1, 2, 4}) {
    const int grid = ncu * wg;
    const double flops = (double)grid * (kThreads / 64) * iters * kAcc * 2.0 * 16 * 16 * 4;
    float ms = 0.0f;
    hipLaunchKernelGGL(mfma_f64_loop, dim3(grid), dim3(kThreads), 0, 0, (double*)out, 10, 1.0);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL(mfma_f64_loop, dim3(grid), dim3(kThreads), 0, 0, (double*)out, iters, 1.0);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("f64 16x16x4, %d workgroup(s) per CU: %.3f ms, %.1f TFLOP/s\n", wg, ms, flops / ms * 1e-9);
    hipLaunchKernelGGL(mfma_f32_loop, dim3(grid), dim3(kThreads), 0, 0, (float*)out, 10, 1.0f);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL(mfma_f32_loop, dim3(grid), dim3(kThreads), 0, 0, (float*)out, iters, 1.0f);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("f32 16x16x4, %d workgroup(s) per CU: %.3f ms, %.1f TFLOP/s\n", wg, ms, flops / ms * 1e-9);
  }
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  CHECK(hipFree(out));
  return 0;
}
