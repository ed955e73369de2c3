// rcg_dense.hpp -- small dense FP64 host routines of the solve:
//  * Cholesky of the b x b Gram matrix for CholQR (replaces cuSOLVER
//    geqrf/orgqr, LMC:1979-2107, by a Gram-based QR);
//  * cyclic-threshold Jacobi for the b x b projected operator (the role of
//    JacobiDiagonalizer.h:251-378 / DSYEV in RayleighChebyshevLM.h:478-527);
//  * Sturm-count bisection for the Lanczos tridiagonal (the role of DSTEBZ,
//    LanczosMaxMinFinder.h:611-723).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>
#include <vector>

namespace rcg {
namespace dense {

// Upper-triangular R with G = R^T R (G column-major n x n, only the upper
// triangle read).  Returns false on a non-positive pivot.
inline bool cholesky_upper(const double* G, long n, std::vector<double>& R) {
  R.assign(n * n, 0.0);
  for (long j = 0; j < n; ++j) {
    double d = G[j + j * n];
    for (long k = 0; k < j; ++k) d -= R[k + j * n] * R[k + j * n];
    if (!(d > 0.0) || !std::isfinite(d)) return false;
    const double rjj = std::sqrt(d);
    R[j + j * n] = rjj;
    for (long c = j + 1; c < n; ++c) {
      double s = G[j + c * n];
      for (long k = 0; k < j; ++k) s -= R[k + j * n] * R[k + c * n];
      R[j + c * n] = s / rjj;
    }
  }
  return true;
}

// Inverse of an upper-triangular matrix (column-major), in a new array.
inline void upper_inverse(const std::vector<double>& R, long n, std::vector<double>& Ri) {
  Ri.assign(n * n, 0.0);
  for (long j = 0; j < n; ++j) {
    Ri[j + j * n] = 1.0 / R[j + j * n];
    for (long i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (long k = i + 1; k <= j; ++k) s += R[i + k * n] * Ri[k + j * n];
      Ri[i + j * n] = -s / R[i + i * n];
    }
  }
}

// Symmetric eigensystem by cyclic Jacobi with an RMS threshold, ascending
// eigenvalues, eigenvectors in the columns of Z (column-major).  H holds the
// full symmetric matrix (column-major); it is not modified.
inline void jacobi_eigensystem(const double* H, long n, std::vector<double>& w,
                               std::vector<double>& Z) {
  std::vector<double> A(H, H + n * n);
  std::vector<double> V(n * n, 0.0);
  for (long i = 0; i < n; ++i) V[i + i * n] = 1.0;
  auto a = [&](long i, long j) -> double& { return A[i + j * n]; };
  for (long j = 0; j < n; ++j)
    for (long i = j + 1; i < n; ++i) a(i, j) = a(j, i) = 0.5 * (a(i, j) + a(j, i));
  double frob = 0.0;
  for (double x : A) frob += x * x;
  frob = std::sqrt(frob);
  const double npairs = n > 1 ? (n * (n - 1)) / 2.0 : 1.0;
  for (int sweep = 0; sweep < 100 && n > 1; ++sweep) {
    double off = 0.0;
    for (long j = 0; j < n; ++j)
      for (long i = 0; i < j; ++i) off += a(i, j) * a(i, j);
    if (off <= (1e-17 * frob) * (1e-17 * frob) || off == 0.0) break;
    // rotate entries above the RMS size first; the final sweeps take all
    const double thr = (sweep < 3) ? std::sqrt(off / npairs) * 0.5 : 0.0;
    for (long p = 0; p < n - 1; ++p) {
      for (long q = p + 1; q < n; ++q) {
        const double apq = a(p, q);
        if (apq == 0.0 || std::abs(apq) < thr) continue;
        const double app = a(p, p), aqq = a(q, q);
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::abs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t);
        const double s = t * c;
        for (long k = 0; k < n; ++k) {
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (long k = 0; k < n; ++k) {
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        a(p, q) = a(q, p) = 0.0;
        for (long k = 0; k < n; ++k) {
          const double vkp = V[k + p * n], vkq = V[k + q * n];
          V[k + p * n] = c * vkp - s * vkq;
          V[k + q * n] = s * vkp + c * vkq;
        }
      }
    }
  }
  std::vector<long> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](long x, long y) { return a(x, x) < a(y, y); });
  w.resize(n);
  Z.assign(n * n, 0.0);
  for (long j = 0; j < n; ++j) {
    w[j] = a(idx[j], idx[j]);
    for (long i = 0; i < n; ++i) Z[i + j * n] = V[i + idx[j] * n];
  }
}

// ---------------------------------------------------------------- complex
using cplx = std::complex<double>;

// Upper-triangular R with G = R^H R (G Hermitian, column-major; upper
// triangle read).  Returns false on a non-positive pivot.
inline bool cholesky_upper(const cplx* G, long n, std::vector<cplx>& R) {
  R.assign(n * n, cplx(0.0));
  for (long j = 0; j < n; ++j) {
    double d = G[j + j * n].real();
    for (long k = 0; k < j; ++k) d -= std::norm(R[k + j * n]);
    if (!(d > 0.0) || !std::isfinite(d)) return false;
    const double rjj = std::sqrt(d);
    R[j + j * n] = rjj;
    for (long c = j + 1; c < n; ++c) {
      cplx s = G[j + c * n];
      for (long k = 0; k < j; ++k) s -= std::conj(R[k + j * n]) * R[k + c * n];
      R[j + c * n] = s / rjj;
    }
  }
  return true;
}

inline void upper_inverse(const std::vector<cplx>& R, long n, std::vector<cplx>& Ri) {
  Ri.assign(n * n, cplx(0.0));
  for (long j = 0; j < n; ++j) {
    Ri[j + j * n] = 1.0 / R[j + j * n];
    for (long i = j - 1; i >= 0; --i) {
      cplx s = 0.0;
      for (long k = i + 1; k <= j; ++k) s += R[i + k * n] * Ri[k + j * n];
      Ri[i + j * n] = -s / R[i + i * n];
    }
  }
}

// Hermitian eigensystem by cyclic Jacobi: each pivot (p, q) first takes the
// phase of H(p, q) into column q (a diagonal unitary), which leaves a real
// symmetric 2x2 block, then a real rotation as in the real routine.
// Ascending eigenvalues w, eigenvectors in the columns of Z.
inline void jacobi_eigensystem(const cplx* H, long n, std::vector<double>& w,
                               std::vector<cplx>& Z) {
  std::vector<cplx> A(H, H + n * n);
  std::vector<cplx> V(n * n, cplx(0.0));
  for (long i = 0; i < n; ++i) V[i + i * n] = 1.0;
  auto a = [&](long i, long j) -> cplx& { return A[i + j * n]; };
  for (long j = 0; j < n; ++j) {
    a(j, j) = a(j, j).real();
    for (long i = j + 1; i < n; ++i) {
      const cplx m = 0.5 * (a(i, j) + std::conj(a(j, i)));
      a(i, j) = m;
      a(j, i) = std::conj(m);
    }
  }
  double frob = 0.0;
  for (const cplx& x : A) frob += std::norm(x);
  frob = std::sqrt(frob);
  const double npairs = n > 1 ? (n * (n - 1)) / 2.0 : 1.0;
  for (int sweep = 0; sweep < 100 && n > 1; ++sweep) {
    double off = 0.0;
    for (long j = 0; j < n; ++j)
      for (long i = 0; i < j; ++i) off += std::norm(a(i, j));
    if (off <= (1e-17 * frob) * (1e-17 * frob) || off == 0.0) break;
    const double thr = (sweep < 3) ? std::sqrt(off / npairs) * 0.5 : 0.0;
    for (long p = 0; p < n - 1; ++p) {
      for (long q = p + 1; q < n; ++q) {
        const double mag = std::abs(a(p, q));
        if (mag == 0.0 || mag < thr) continue;
        const cplx d = std::conj(a(p, q)) / mag;  // column q *= d, row q *= conj(d)
        for (long k = 0; k < n; ++k) a(k, q) *= d;
        for (long k = 0; k < n; ++k) a(q, k) *= std::conj(d);
        for (long k = 0; k < n; ++k) V[k + q * n] *= d;
        const double app = a(p, p).real(), aqq = a(q, q).real();
        const double tau = (aqq - app) / (2.0 * mag);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::abs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t);
        const double s = t * c;
        for (long k = 0; k < n; ++k) {
          const cplx akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (long k = 0; k < n; ++k) {
          const cplx apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        a(p, q) = a(q, p) = 0.0;
        a(p, p) = a(p, p).real();
        a(q, q) = a(q, q).real();
        for (long k = 0; k < n; ++k) {
          const cplx vkp = V[k + p * n], vkq = V[k + q * n];
          V[k + p * n] = c * vkp - s * vkq;
          V[k + q * n] = s * vkp + c * vkq;
        }
      }
    }
  }
  std::vector<long> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(),
                   [&](long x, long y) { return a(x, x).real() < a(y, y).real(); });
  w.resize(n);
  Z.assign(n * n, cplx(0.0));
  for (long j = 0; j < n; ++j) {
    w[j] = a(idx[j], idx[j]).real();
    for (long i = 0; i < n; ++i) Z[i + j * n] = V[i + idx[j] * n];
  }
}

// idx-th smallest eigenvalue of the symmetric tridiagonal (d, e) by Sturm
// bisection to width max(abstol, 2 eps |end|).
inline double tridiag_eigenvalue(const std::vector<double>& d, const std::vector<double>& e,
                                 long idx, double abstol) {
  const long n = (long)d.size();
  double lo = d[0], hi = d[0], e2max = 0.0;
  for (long i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::abs(e[i - 1]) : 0.0) + (i + 1 < n ? std::abs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
    if (i + 1 < n) e2max = std::max(e2max, e[i] * e[i]);
  }
  const double tiny = 2.2250738585072014e-308;
  const double eps = 2.220446049250313e-16;
  const double pivmin = std::max(tiny, tiny * e2max);
  const double pad = 2.0 * eps * std::max(std::abs(lo), std::abs(hi)) + 2.0 * pivmin;
  lo -= pad;
  hi += pad;
  auto below = [&](double x) {
    long c = 0;
    double q = d[0] - x;
    if (std::abs(q) < pivmin) q = -pivmin;
    if (q < 0) ++c;
    for (long i = 1; i < n; ++i) {
      q = d[i] - x - (e[i - 1] * e[i - 1]) / q;
      if (std::abs(q) < pivmin) q = -pivmin;
      if (q < 0) ++c;
    }
    return c;
  };
  for (int it = 0; it < 200; ++it) {
    if (hi - lo <= std::max(abstol, 2.0 * eps * std::max(std::abs(lo), std::abs(hi)))) break;
    const double mid = 0.5 * (lo + hi);
    if (below(mid) >= idx + 1) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

}  // namespace dense
}  // namespace rcg
