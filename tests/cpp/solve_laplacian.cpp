// Drop-in check of include/rcg/RayleighChebyshevLM.hpp: the calling sequence
// of the reference's getMinEigenSystem (RayleighChebyshevLM.h:712-738) on a
// 2D Dirichlet Laplacian whose eigenvalues are known in closed form.
//
//   usage: solve_laplacian <nx> <k> [tol]
// prints one line per eigenpair "lambda exact residual" and exits non-zero
// when an eigenvalue or residual misses the tolerance.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "rcg/RayleighChebyshevLM.hpp"

int main(int argc, char** argv) {
  const long nx = argc > 1 ? std::atol(argv[1]) : 32;
  const long k = argc > 2 ? std::atol(argv[2]) : 8;
  const double tol = argc > 3 ? std::atof(argv[3]) : 1e-10;
  const long n = nx * nx;
  std::vector<int32_t> rowptr{0}, col;
  std::vector<double> val;
  for (long y = 0; y < nx; ++y)
    for (long x = 0; x < nx; ++x) {
      const long r = x + nx * y;
      const long nb[4] = {x > 0 ? r - 1 : -1, x + 1 < nx ? r + 1 : -1, y > 0 ? r - nx : -1,
                          y + 1 < nx ? r + nx : -1};
      col.push_back((int32_t)r);
      val.push_back(4.0);
      for (long c : nb)
        if (c >= 0) {
          col.push_back((int32_t)c);
          val.push_back(-1.0);
        }
      rowptr.push_back((int32_t)col.size());
    }

  rcg::Context ctx(0);
  rcg::CsrOperator op(ctx, n, rowptr, col, val);
  op.setGrid(nx, nx, 1);
  rcg::AvectorClass<double> vStart(n);
  rcg::ArandomizeOpClass randOp(7);
  rcg::AmatrixClass<double> eigVectors(ctx, n, k);
  std::vector<double> eigValues;

  rcg::RayleighChebyshevLM<rcg::AmatrixClass<double>, rcg::AvectorClass<double>, rcg::CsrOperator,
                           rcg::ArandomizeOpClass>
      rc;
  rc.setStopCondition("RESIDUAL_ONLY");
  const long found = rc.getMinEigenSystem(k, tol, 8, 8, vStart, op, randOp, eigValues, eigVectors);

  std::vector<double> exact;
  const double pi = std::acos(-1.0);
  for (long a = 1; a <= nx; ++a)
    for (long b = 1; b <= nx; ++b)
      exact.push_back(4.0 - 2.0 * std::cos(pi * a / (nx + 1)) - 2.0 * std::cos(pi * b / (nx + 1)));
  std::sort(exact.begin(), exact.end());

  // Residuals recomputed on the host from the downloaded eigenvectors.
  eigVectors.device_to_host();
  const std::vector<double> res = rc.getEigVectorResiduals();
  int bad = found < k;
  for (long j = 0; j < found; ++j) {
    double rn = 0, vn = 0;
    for (long r = 0; r < n; ++r) {
      double av = 0;
      for (int32_t p = rowptr[r]; p < rowptr[r + 1]; ++p) av += val[p] * eigVectors(col[p], j);
      const double d = av - eigValues[j] * eigVectors(r, j);
      rn += d * d;
      vn += eigVectors(r, j) * eigVectors(r, j);
    }
    const double resid = std::sqrt(rn) / 8.0;  // ||A|| <= 8 for this operator
    const double err = std::fabs(eigValues[j] - exact[j]);
    std::printf("%.15f %.15f %.3e %.3e %.3e\n", eigValues[j], exact[j], res[j], resid,
                std::fabs(vn - 1.0));
    if (err > 1e-10 * std::max(1.0, std::fabs(exact[j])) || resid > 10 * tol ||
        std::fabs(vn - 1.0) > 1e-10)
      bad = 1;
  }
  // The CPU template's form: eigenvectors as host vectors, same seed and
  // stream, so the same eigenvalues to round-off.
  rcg::ArandomizeOpClass randOp2(7);
  std::vector<double> eigValues2;
  std::vector<rcg::AvectorClass<double>> eigVectors2;
  const long found2 =
      rc.getMinEigenSystem(k, tol, 8, 8, vStart, op, randOp2, eigValues2, eigVectors2);
  if (found2 != found || (long)eigVectors2.size() != found2) bad = 1;
  for (long j = 0; j < std::min(found, found2); ++j) {
    const double nrm = eigVectors2[j].norm2();
    if (std::fabs(eigValues2[j] - eigValues[j]) > 1e-12 * std::max(1.0, std::fabs(eigValues[j])) ||
        std::fabs(nrm - 1.0) > 1e-10)
      bad = 1;
  }
  // The older RayleighChebyshev template: host vectors, and an interval solve
  // with the caller's bounds (no Lanczos) for every eigenvalue below a gap.
  rcg::RayleighChebyshev<rcg::AvectorClass<double>, rcg::CsrOperator, rcg::ArandomizeOpClass> rc3;
  rc3.setStopCondition("RESIDUAL_ONLY");
  long below = k;
  while (below + 1 < (long)exact.size() && exact[below + 1] - exact[below] < 1e-6) ++below;
  const double lambdaMax = 0.5 * (exact[below] + exact[below + 1]);
  rcg::ArandomizeOpClass randOp3(3);
  std::vector<double> eigValues3;
  std::vector<rcg::AvectorClass<double>> eigVectors3;
  const long found3 = rc3.getMinIntervalEigenSystem(0.0, lambdaMax, 8.0, tol, 8, 8, 4 * k, vStart,
                                                    op, randOp3, eigValues3, eigVectors3);
  if (found3 != below + 1 || (long)eigVectors3.size() != found3 ||
      rc3.getSpectralBoundsIterCount() != 0)
    bad = 1;
  for (long j = 0; j < std::min(found3, below + 1); ++j)
    if (std::fabs(eigValues3[j] - exact[j]) > 1e-10 * std::max(1.0, exact[j])) bad = 1;
  std::printf("interval below %.6f: found %ld of %ld\n", lambdaMax, found3, below + 1);
  const auto cd = rc.getCountData();
  std::printf("found %ld ortho %ld apply %ld eig %ld lanczos %ld\n", found, cd.at("ortho"),
              cd.at("OpApply"), cd.at("eigenvalue"), rc.getSpectralBoundsIterCount());
  return bad;
}
