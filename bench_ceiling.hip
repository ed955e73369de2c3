// bench_ceiling.hip -- achievable-rate kernels for the FP64 Chebyshev step's
// memory pattern, used by bench.py to put the step's measured rate next to
// what the same accesses reach without the CSR operator.
//
// The step (csr_block_kernel, kCheb) on an n x b FP64 block, row-interleaved
// (element (i, c) at X[i*b + c]): 16 lanes own a row of b = 64 columns, two
// 16-byte vectors per lane; the row's X_{j-1} entries are read, its X_{j-2}
// entries are read once and its X_j entries written once (non-temporal), and
// every nonzero of the row gathers a whole row of X_{j-1}.  Two variants keep
// exactly those row accesses:
//   kind 0 "stream":  out = p2 (p1 X - Z), no gathers (two reads, one write
//                     of the block: 3 n b 8 bytes);
//   kind 1 "stencil": out = p2 (p1 (p0 A X + X) - Z) with A the periodic
//                     7-point Laplacian of an nx^3 grid computed from the row
//                     index (the step's gathers without the CSR arrays).
// ceiling_ms(kind, nx, iters) allocates the blocks, times `iters` launches with
// HIP events and returns the mean milliseconds per launch (negative on error).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kCols = 64;                 // b
constexpr int kLanes = 16;                // lanes per row
constexpr int kVec = 2;                   // 16-byte vectors per lane
constexpr int kRowsPerPass = 256 / kLanes;
constexpr int kRowsPerBlock = 128;

typedef double dvec2 __attribute__((ext_vector_type(2)));

template <int KIND>
__global__ __launch_bounds__(256) void ceiling_kernel(int64_t nx, const double* __restrict__ X,
                                                      const double* __restrict__ Z,
                                                      double* __restrict__ out, double p0,
                                                      double p1, double p2) {
  const int64_t n = nx * nx * nx;
  const int g = threadIdx.x % kLanes;
  const int64_t i0 = (int64_t)blockIdx.x * kRowsPerBlock;
  for (int r = threadIdx.x / kLanes; r < kRowsPerBlock; r += kRowsPerPass) {
    const int64_t i = i0 + r;
    if (i >= n) break;
    dvec2 x[kVec], z[kVec], acc[kVec];
#pragma unroll
    for (int k = 0; k < kVec; ++k) {
      const int64_t off = i * kCols + 2 * (g + k * kLanes);
      x[k] = *reinterpret_cast<const dvec2*>(X + off);
      z[k] = __builtin_nontemporal_load(reinterpret_cast<const dvec2*>(Z + off));
      acc[k] = dvec2{0.0, 0.0};
    }
    if constexpr (KIND == 1) {
      const int64_t xx = i % nx, yy = (i / nx) % nx, zz = i / (nx * nx);
      int64_t nb[7];
      nb[0] = xx + nx * (yy + nx * ((zz + nx - 1) % nx));
      nb[1] = xx + nx * (((yy + nx - 1) % nx) + nx * zz);
      nb[2] = (xx + nx - 1) % nx + nx * (yy + nx * zz);
      nb[3] = i;
      nb[4] = (xx + 1) % nx + nx * (yy + nx * zz);
      nb[5] = xx + nx * (((yy + 1) % nx) + nx * zz);
      nb[6] = xx + nx * (yy + nx * ((zz + 1) % nx));
      // chunks of 4 gathers issued before their FMAs, as in the step
      for (int c = 0; c < 7; c += 4) {
        dvec2 u[4][kVec];
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (c + e < 7)
#pragma unroll
            for (int k = 0; k < kVec; ++k)
              u[e][k] = *reinterpret_cast<const dvec2*>(X + nb[c + e] * kCols + 2 * (g + k * kLanes));
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (c + e < 7) {
            const double a = (c + e == 3) ? 6.0 : -1.0;
#pragma unroll
            for (int k = 0; k < kVec; ++k) {
              acc[k].x = fma(a, u[e][k].x, acc[k].x);
              acc[k].y = fma(a, u[e][k].y, acc[k].y);
            }
          }
      }
    }
#pragma unroll
    for (int k = 0; k < kVec; ++k) {
      dvec2 o;
      o.x = p2 * fma(p1, fma(p0, acc[k].x, x[k].x), -z[k].x);
      o.y = p2 * fma(p1, fma(p0, acc[k].y, x[k].y), -z[k].y);
      __builtin_nontemporal_store(o, reinterpret_cast<dvec2*>(out + i * kCols + 2 * (g + k * kLanes)));
    }
  }
}

__global__ void fill_kernel(double* p, int64_t count, double v) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count;
       t += (int64_t)gridDim.x * blockDim.x)
    p[t] = v + 1e-9 * (double)(t % 1021);
}

}  // namespace

extern "C" double ceiling_ms(int kind, int64_t nx, int iters) {
  if ((kind != 0 && kind != 1) || nx < 3 || iters < 1) return -1.0;
  const int64_t n = nx * nx * nx;
  const size_t bytes = sizeof(double) * (size_t)n * kCols;
  double *X = nullptr, *Z = nullptr, *O = nullptr;
  if (hipMalloc(&X, bytes) != hipSuccess) return -2.0;
  if (hipMalloc(&Z, bytes) != hipSuccess) {
    (void)hipFree(X);
    return -2.0;
  }
  if (hipMalloc(&O, bytes) != hipSuccess) {
    (void)hipFree(X);
    (void)hipFree(Z);
    return -2.0;
  }
  const int64_t count = n * kCols;
  hipLaunchKernelGGL(fill_kernel, dim3(4096), dim3(256), 0, 0, X, count, 1.0);
  hipLaunchKernelGGL(fill_kernel, dim3(4096), dim3(256), 0, 0, Z, count, 0.5);
  const dim3 grid((unsigned)((n + kRowsPerBlock - 1) / kRowsPerBlock));
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  double ms = -3.0;
  for (int rep = 0; rep < 2; ++rep) {  // the first round warms up
    (void)hipEventRecord(e0, 0);
    for (int it = 0; it < iters; ++it) {
      if (kind == 0)
        hipLaunchKernelGGL(ceiling_kernel<0>, grid, dim3(256), 0, 0, nx, X, Z, O, 0.1, 0.9, 1.0);
      else
        hipLaunchKernelGGL(ceiling_kernel<1>, grid, dim3(256), 0, 0, nx, X, Z, O, 0.1, 0.9, 1.0);
    }
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) break;
    float t = 0.f;
    if (hipEventElapsedTime(&t, e0, e1) == hipSuccess) ms = t / iters;
  }
  if (hipGetLastError() != hipSuccess) ms = -4.0;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(X);
  (void)hipFree(Z);
  (void)hipFree(O);
  return ms;
}
