// Matrix-free operators for the caller-operator tests: compiled applies with
// the rcg_apply_fn signature of include/rcg_c_api.h, the way a caller of the
// reference writes its Otype::apply(in, out, alpha, beta)
// (LanczosCpolyOperatorLMCuda.h:236,245).  No matrix is stored: the 3-D
// periodic 7-point Laplacian plus a diagonal potential, and the 2-D periodic
// Hofstadter lattice of operators.hofstadter_2d (hopping -1 on x-bonds,
// -exp(2 pi i phi x) on the hop (x, y) -> (x, y+1)), on column-major blocks.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace {

struct Anderson3d {
  int64_t L;
  double* diag;  // device: 6 + V_i
};

struct Hofstadter {
  int64_t L;
  double phi;
  double* onsite;  // device
};

__global__ void anderson3d_kernel(int64_t L, const double* diag, int64_t cols, int64_t ld,
                                  const double* in, double* out, double alpha, double beta) {
  const int64_t n = L * L * L;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * cols) return;
  const int64_t c = t / n, i = t % n;
  const int64_t x = i % L, y = (i / L) % L, z = i / (L * L);
  const double* u = in + c * ld;
  auto at = [&](int64_t xx, int64_t yy, int64_t zz) {
    return u[((xx + L) % L) + L * (((yy + L) % L) + L * ((zz + L) % L))];
  };
  double a = diag[i] * u[i];
  a -= at(x - 1, y, z) + at(x + 1, y, z);
  a -= at(x, y - 1, z) + at(x, y + 1, z);
  a -= at(x, y, z - 1) + at(x, y, z + 1);
  double* o = out + c * ld + i;
  *o = beta == 0.0 ? alpha * a : alpha * a + beta * *o;
}

__global__ void hofstadter_kernel(int64_t L, double phi, const double* onsite, int64_t cols,
                                  int64_t ld, const double2* in, double2* out, double alpha,
                                  double beta) {
  const int64_t n = L * L;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * cols) return;
  const int64_t c = t / n, i = t % n;
  const int64_t x = i % L, y = i / L;
  const double2* u = in + c * ld;
  const double th = 2.0 * M_PI * phi * (double)x;
  const double cr = cos(th), ci = sin(th);  // ph(x) = exp(i th)
  const double2 xm = u[((x - 1 + L) % L) + L * y], xp = u[((x + 1) % L) + L * y];
  const double2 yp = u[x + L * ((y + 1) % L)], ym = u[x + L * ((y - 1 + L) % L)];
  // A u = onsite u - u(x-1) - u(x+1) - conj(ph) u(y+1) - ph u(y-1)
  double ar = onsite[i] * u[i].x - xm.x - xp.x;
  double ai = onsite[i] * u[i].y - xm.y - xp.y;
  ar -= cr * yp.x + ci * yp.y;
  ai -= cr * yp.y - ci * yp.x;
  ar -= cr * ym.x - ci * ym.y;
  ai -= cr * ym.y + ci * ym.x;
  double2* o = out + c * ld + i;
  if (beta == 0.0) {
    *o = make_double2(alpha * ar, alpha * ai);
  } else {
    const double2 p = *o;
    *o = make_double2(alpha * ar + beta * p.x, alpha * ai + beta * p.y);
  }
}

unsigned blocks_for(int64_t work) { return (unsigned)((work + 255) / 256); }

}  // namespace

extern "C" {

void* anderson3d_create(int64_t L, const double* diag_host) {
  auto* p = new Anderson3d{L, nullptr};
  if (hipMalloc(&p->diag, sizeof(double) * L * L * L) != hipSuccess ||
      hipMemcpy(p->diag, diag_host, sizeof(double) * L * L * L, hipMemcpyHostToDevice) != hipSuccess) {
    delete p;
    return nullptr;
  }
  return p;
}

void* hofstadter_create(int64_t L, double phi, const double* onsite_host) {
  auto* p = new Hofstadter{L, phi, nullptr};
  if (hipMalloc(&p->onsite, sizeof(double) * L * L) != hipSuccess ||
      hipMemcpy(p->onsite, onsite_host, sizeof(double) * L * L, hipMemcpyHostToDevice) != hipSuccess) {
    delete p;
    return nullptr;
  }
  return p;
}

void anderson3d_destroy(void* user) {
  auto* p = static_cast<Anderson3d*>(user);
  (void)hipFree(p->diag);
  delete p;
}

void hofstadter_destroy(void* user) {
  auto* p = static_cast<Hofstadter*>(user);
  (void)hipFree(p->onsite);
  delete p;
}

int anderson3d_apply(void* user, const void* in, void* out, int64_t n, int64_t cols, int64_t ld,
                     double alpha, double beta, void* stream) {
  const auto* p = static_cast<const Anderson3d*>(user);
  if (n != p->L * p->L * p->L) return 2;
  hipLaunchKernelGGL(anderson3d_kernel, dim3(blocks_for(n * cols)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), p->L, p->diag, cols, ld,
                     static_cast<const double*>(in), static_cast<double*>(out), alpha, beta);
  return hipGetLastError() == hipSuccess ? 0 : 3;
}

int hofstadter_apply(void* user, const void* in, void* out, int64_t n, int64_t cols, int64_t ld,
                     double alpha, double beta, void* stream) {
  const auto* p = static_cast<const Hofstadter*>(user);
  if (n != p->L * p->L) return 2;
  hipLaunchKernelGGL(hofstadter_kernel, dim3(blocks_for(n * cols)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), p->L, p->phi, p->onsite, cols, ld,
                     static_cast<const double2*>(in), static_cast<double2*>(out), alpha, beta);
  return hipGetLastError() == hipSuccess ? 0 : 3;
}

}  // extern "C"
