// Drop-in check of the caller-operator adapter of
// include/rcg/RayleighChebyshevLM.hpp: a matrix-free operator written the way
// a reference caller writes its Otype (apply(in, out, alpha, beta) on device
// blocks, LanczosCpolyOperatorLMCuda.h:236,245) -- here a HIP kernel for the
// 2-D Dirichlet 5-point Laplacian on column-major blocks -- solved through
// getMinEigenSystem; eigenvalues are known in closed form.
//
//   usage: solve_matrix_free <nx> <k> [tol]
// prints one line per eigenpair "lambda exact" and exits non-zero when an
// eigenvalue misses the tolerance or the residual check fails.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "rcg/RayleighChebyshevLM.hpp"

__global__ void laplace2d_kernel(long nx, long cols, long ld, const double* in, double* out,
                                 double alpha, double beta) {
  const long n = nx * nx;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * cols) return;
  const long c = t / n, i = t % n, x = i % nx, y = i / nx;
  const double* u = in + c * ld;
  double a = 4.0 * u[i];
  if (x > 0) a -= u[i - 1];
  if (x + 1 < nx) a -= u[i + 1];
  if (y > 0) a -= u[i - nx];
  if (y + 1 < nx) a -= u[i + nx];
  double* o = out + c * ld + i;
  *o = beta == 0.0 ? alpha * a : alpha * a + beta * *o;
}

// The caller's operator: no matrix, only the stencil.
struct Laplace2D {
  long nx;
  long applies = 0;
  void apply(rcg::DeviceMatrix<double>& in, rcg::DeviceMatrix<double>& out, double alpha,
             double beta) {
    const long work = in.getRowSize() * in.getColSize();
    hipLaunchKernelGGL(laplace2d_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(in.stream()), nx, in.getColSize(),
                       in.getLeadingDimension(), in.getDataPointer_d(), out.getDataPointer_d(),
                       alpha, beta);
    if (hipGetLastError() != hipSuccess) throw std::runtime_error("stencil launch failed");
    ++applies;
  }
};

int main(int argc, char** argv) {
  const long nx = argc > 1 ? std::atol(argv[1]) : 32;
  const long k = argc > 2 ? std::atol(argv[2]) : 8;
  const double tol = argc > 3 ? std::atof(argv[3]) : 1e-10;
  const long n = nx * nx;

  rcg::Context ctx(0);
  Laplace2D stencil{nx};
  rcg::CallbackOperator<Laplace2D> op(ctx, n, stencil);
  rcg::AvectorClass<double> vStart(n);
  rcg::ArandomizeOpClass randOp(7);
  rcg::AmatrixClass<double> eigVectors(ctx, n, k);
  std::vector<double> eigValues;
  rcg::RayleighChebyshevLM<rcg::AmatrixClass<double>, rcg::AvectorClass<double>,
                           rcg::CallbackOperator<Laplace2D>, rcg::ArandomizeOpClass>
      rc;
  rc.setStopCondition("RESIDUAL_ONLY");
  const long found = rc.getMinEigenSystem(k, tol, 8, 8, vStart, op, randOp, eigValues, eigVectors);

  std::vector<double> exact;
  const double pi = std::acos(-1.0);
  for (long a = 1; a <= nx; ++a)
    for (long b = 1; b <= nx; ++b)
      exact.push_back(4.0 - 2.0 * std::cos(pi * a / (nx + 1)) - 2.0 * std::cos(pi * b / (nx + 1)));
  std::sort(exact.begin(), exact.end());

  // residuals through the same caller operator: W = A V, r_j = ||W_j - lambda_j V_j||
  rcg::AmatrixClass<double> W(ctx, n, eigVectors.getColSize());
  op.apply(eigVectors, W, 1.0, 0.0);
  eigVectors.device_to_host();
  W.device_to_host();
  int bad = found < k || stencil.applies == 0;
  for (long j = 0; j < found; ++j) {
    double rn = 0;
    for (long r = 0; r < n; ++r) {
      const double d = W(r, j) - eigValues[j] * eigVectors(r, j);
      rn += d * d;
    }
    const double err = std::fabs(eigValues[j] - exact[j]);
    std::printf("%.15f %.15f %.3e\n", eigValues[j], exact[j], std::sqrt(rn) / 8.0);
    if (err > 1e-10 * std::max(1.0, std::fabs(exact[j])) || std::sqrt(rn) / 8.0 > 10 * tol) bad = 1;
  }
  std::printf("found %ld applies %ld lanczos %ld\n", found, stencil.applies,
              rc.getSpectralBoundsIterCount());
  return bad;
}
