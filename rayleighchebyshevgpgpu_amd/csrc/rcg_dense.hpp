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
#include <type_traits>
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

// ------------------------------------------- Householder + implicit QL
// Eigensystem of a symmetric (real) or Hermitian (complex) n x n matrix in
// O(n^3) with a small constant, for the larger Rayleigh-Ritz problems:
//  1. Householder reduction A = Q T Q^H with T real symmetric tridiagonal
//     (for complex A the off-diagonal phases are moved into Q);
//  2. implicit QL iteration with Wilkinson-type shifts on T (the tql2 method
//     of Bowdler, Martin, Reinsch and Wilkinson, 1968), each plane rotation
//     applied to two contiguous columns of Q.
// Ascending eigenvalues w, eigenvectors in the columns of Z.
inline double conj_v(double x) { return x; }
inline cplx conj_v(cplx x) { return std::conj(x); }
inline double real_v(double x) { return x; }
inline double real_v(cplx x) { return x.real(); }

// sum_i conj(x_i) y_i with eight partial sums, which the compiler keeps in
// vector registers (a single running sum would serialise the loop).
template <class S>
inline S dotc(const S* x, const S* y, long m) {
  S p[8] = {S(0.0), S(0.0), S(0.0), S(0.0), S(0.0), S(0.0), S(0.0), S(0.0)};
  long i = 0;
  for (; i + 8 <= m; i += 8)
    for (int u = 0; u < 8; ++u) p[u] += conj_v(x[i + u]) * y[i + u];
  for (; i < m; ++i) p[i & 7] += conj_v(x[i]) * y[i];
  return ((p[0] + p[4]) + (p[1] + p[5])) + ((p[2] + p[6]) + (p[3] + p[7]));
}

// Only the lower triangle of the trailing block is read and updated: the
// block stays Hermitian, the diagonal, the subdiagonal and every Householder
// column lie in the lower triangle.
template <class S>
inline void tridiagonalize(std::vector<S>& A, long n, std::vector<double>& d,
                           std::vector<double>& e, std::vector<S>& Q) {
  auto a = [&](long i, long j) -> S& { return A[i + j * n]; };
  d.assign(n, 0.0);
  e.assign(n, 0.0);
  std::vector<std::vector<S>> vs(n > 2 ? n - 2 : 0);
  std::vector<double> betas(n > 2 ? n - 2 : 0, 0.0);
  std::vector<S> sub(n, S(0.0));  // complex subdiagonal before the phase fix
  std::vector<S> p(n), w(n);
  for (long k = 0; k + 2 < n; ++k) {
    const long m = n - k - 1;
    double sx = 0.0;
    for (long i = 1; i < m; ++i) sx += std::norm(a(k + 1 + i, k));
    const S x0 = a(k + 1, k);
    if (sx == 0.0) {  // column already tridiagonal
      sub[k] = x0;
      continue;
    }
    const double nx = std::sqrt(sx + std::norm(x0));
    const double ax0 = std::abs(x0);
    const S ph = ax0 > 0.0 ? x0 / ax0 : S(1.0);
    const S alpha = -ph * nx;
    std::vector<S>& v = vs[k];
    v.assign(m, S(0.0));
    v[0] = x0 - alpha;
    for (long i = 1; i < m; ++i) v[i] = a(k + 1 + i, k);
    const double vv = real_v(dotc(v.data(), v.data(), m));
    const double beta = 2.0 / vv;
    betas[k] = beta;
    sub[k] = alpha;
    // trailing block A22 = A[k+1.., k+1..]: p = beta A22 v, w = p - K v,
    // K = beta (v^H p) / 2, A22 -= v w^H + w v^H.  Column j of the lower
    // triangle gives p[i > j] += A(i, j) v_j (contiguous axpy) and, through
    // A(j, i) = conj(A(i, j)), p[j] += sum_{i > j} conj(A(i, j)) v_i.
    std::fill(p.begin(), p.begin() + m, S(0.0));
    for (long j = 0; j < m; ++j) {
      const S* col = &a(k + 1, k + 1 + j);
      const S vj = v[j];
      for (long i = j + 1; i < m; ++i) p[i] += col[i] * vj;
      p[j] += real_v(col[j]) * vj + dotc(col + j + 1, v.data() + j + 1, m - j - 1);
    }
    for (long i = 0; i < m; ++i) p[i] *= beta;
    const double K = 0.5 * beta * real_v(dotc(v.data(), p.data(), m));
    for (long i = 0; i < m; ++i) w[i] = p[i] - K * v[i];
    for (long j = 0; j < m; ++j) {
      const S cvj = conj_v(v[j]), cwj = conj_v(w[j]);
      S* col = &a(k + 1, k + 1 + j);
      for (long i = j; i < m; ++i) col[i] -= v[i] * cwj + w[i] * cvj;
    }
  }
  for (long k = 0; k < n; ++k) d[k] = real_v(a(k, k));
  if (n >= 2) sub[n - 2] = a(n - 1, n - 2);
  // Q = H_0 H_1 ... H_{n-3}, accumulated from the right end
  Q.assign(n * n, S(0.0));
  for (long i = 0; i < n; ++i) Q[i + i * n] = S(1.0);
  for (long k = n - 3; k >= 0; --k) {
    if (vs[k].empty()) continue;
    const std::vector<S>& v = vs[k];
    const long m = n - k - 1;
    for (long j = k + 1; j < n; ++j) {
      S* qj = &Q[(k + 1) + j * n];
      const S t = betas[k] * dotc(v.data(), qj, m);
      for (long i = 0; i < m; ++i) qj[i] -= v[i] * t;
    }
  }
  // real off-diagonals: T' = D^H T D with D = diag(phases), Q <- Q D
  S phase = S(1.0);
  for (long k = 0; k + 1 < n; ++k) {
    const double mag = std::abs(sub[k]);
    e[k] = mag;
    // column k+1 of Q takes the accumulated phase making sub[k] real and >= 0
    const S t = mag > 0.0 ? sub[k] / mag : S(1.0);
    phase = phase * t;
    if constexpr (std::is_same<S, cplx>::value) {
      for (long i = 0; i < n; ++i) Q[i + (k + 1) * n] *= phase;
    } else {
      if (phase < 0.0)
        for (long i = 0; i < n; ++i) Q[i + (k + 1) * n] = -Q[i + (k + 1) * n];
    }
  }
}

template <class S>
inline void tridiag_ql(std::vector<double>& d, std::vector<double>& e, long n, std::vector<S>& Z) {
  const double eps = 2.220446049250313e-16;
  for (long l = 0; l < n; ++l) {
    int iter = 0;
    long m;
    do {
      for (m = l; m + 1 < n; ++m) {
        const double dd = std::abs(d[m]) + std::abs(d[m + 1]);
        if (std::abs(e[m]) <= eps * dd) break;
      }
      if (m != l) {
        if (++iter > 200) break;  // no convergence: keep what we have
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
        double s = 1.0, c = 1.0, pp = 0.0;
        long i;
        bool early = false;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i];
          const double b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= pp;
            e[m] = 0.0;
            early = true;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - pp;
          r = (d[i] - g) * s + 2.0 * c * b;
          pp = s * r;
          d[i + 1] = g + pp;
          g = c * r - b;
          // real rotation: a complex column is rotated as 2n doubles
          double* zi = reinterpret_cast<double*>(&Z[i * n]);
          double* zj = reinterpret_cast<double*>(&Z[(i + 1) * n]);
          const long len = n * (long)(sizeof(S) / sizeof(double));
          for (long k = 0; k < len; ++k) {
            const double fz = zj[k];
            zj[k] = s * zi[k] + c * fz;
            zi[k] = c * zi[k] - s * fz;
          }
        }
        if (early) continue;
        d[l] -= pp;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
}

template <class S>
inline void tridiag_eigensystem(const S* H, long n, std::vector<double>& w, std::vector<S>& Z) {
  std::vector<S> A(H, H + n * n);
  for (long j = 0; j < n; ++j) {
    A[j + j * n] = real_v(A[j + j * n]);
    for (long i = j + 1; i < n; ++i) {
      const S m = 0.5 * (A[i + j * n] + conj_v(A[j + i * n]));
      A[i + j * n] = m;
      A[j + i * n] = conj_v(m);
    }
  }
  std::vector<double> d, e;
  std::vector<S> Q;
  tridiagonalize(A, n, d, e, Q);
  tridiag_ql(d, e, n, Q);
  std::vector<long> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](long x, long y) { return d[x] < d[y]; });
  w.resize(n);
  Z.assign(n * n, S(0.0));
  for (long j = 0; j < n; ++j) {
    w[j] = d[idx[j]];
    std::copy(Q.begin() + idx[j] * n, Q.begin() + (idx[j] + 1) * n, Z.begin() + j * n);
  }
}

// The Rayleigh-Ritz eigensolver: cyclic Jacobi for small matrices (the
// reference's JacobiDiagonalizer role), Householder + QL from 48 on, where
// Jacobi's O(sweeps n^3) with rotations of whole rows dominates the solve.
template <class S>
inline void small_eigensystem(const S* H, long n, std::vector<double>& w, std::vector<S>& Z) {
  if (n >= 48) tridiag_eigensystem(H, n, w, Z);
  else jacobi_eigensystem(H, n, w, Z);
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
