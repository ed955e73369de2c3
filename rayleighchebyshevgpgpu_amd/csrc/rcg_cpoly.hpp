// rcg_cpoly.hpp -- host scalars of the scaled Lanczos "C" filter polynomial.
//
// The degree/bound choice of the solve follows LanczosCpoly.h exactly
// (evaluatePm 234-270, evaluatePmPrime 400-415, getPmPrimeFirstRoot 371-398,
// getPmNextToLastRoot 363-369, evaluateInversePm 416-473, getXStar 285-293,
// getUpperXStar 295-305, getStarDegreeAndSpectralRadius 498-555).  The
// secant starting points, tolerances and the two pi constants are those of
// the reference, and floating-point operations are kept in the same order,
// so the integer degree sequence of a solve matches the reference's.
#pragma once

#include <cmath>

namespace rcg {
namespace cpoly {

// P_m(x) = b_m/(m+1), b_0 = 1, b_1 = (2-4x) b_0, b_j = (2-4x) b_{j-1} - b_{j-2}.
inline double pm(long m, double x) {
  if (m == 0) return 1.0;
  if (m == 1) return 1.0 - 2.0 * x;
  double prev = 1.0;
  double cur = ((prev * x) * -2.0 + prev) * 2.0;
  for (long k = 2; k <= m; ++k) {
    double nxt = ((cur * x) * -2.0 + cur) * 2.0;
    prev = -prev;
    nxt += prev;
    prev = cur;
    cur = nxt;
  }
  return cur * (1.0 / double(m + 1));
}

// d/dx of sin((m+1)t)/((m+1) sin t) with sin^2(t/2) = x.
inline double pm_prime(long m, double x) {
  const double m1 = double(m) + 1.0;
  const double sx = std::sqrt(x);
  const double t = std::asin(sx);
  const double mt = m1 * t;
  const double isx = 1 / sx;
  const double icx = 1 / std::sqrt(1.0 - x);
  const double s2t = std::sin(2.0 * t);
  return std::cos(2.0 * mt) * isx * icx / s2t -
         std::sin(2.0 * mt) / m1 / (s2t * s2t) * std::cos(2.0 * t) * isx * icx;
}

inline double pm_prime_first_root(long m) {
  const double tol = 1.0e-12;
  const double md = double(m);
  double xa = ((4.49 * 4.49) / (4.0 * (md + 1) * (md + 1)));
  double fa = pm_prime(m, xa);
  double xb = ((4.49 * 4.49) / (4.0 * md * (md + 1)));
  double fb = pm_prime(m, xb);
  for (int it = 0; std::abs(xa - xb) > tol && it < 10000; ++it) {
    const double slope = (fb - fa) / (xb - xa);
    const double xc = xb - pm_prime(m, xb) / slope;
    xa = xb;
    fa = fb;
    xb = xc;
    fb = pm_prime(m, xb);
  }
  return xb;
}

inline double pm_next_to_last_root(long m) {
  const double pi = 3.1415926535897932384;
  return (std::cos((2.0 / (double(m) + 1.0)) * pi) + 1) * 0.5;
}

inline double pm_inverse(long m, double y) {
  const double tol = 1.0e-12;
  const double md = double(m);
  const double pi_r = 3.14159265359;
  if (y < 0.0) {
    const double s = std::sin(0.5 * (pi_r / (md + 1)));
    return s * s;
  }
  double xa = 0.0, fa = 0.0, xb = 0.0, fb = 0.0;
  if (y <= 5.0) {
    xa = 0.0;
    fa = 1.0;
    xb = std::sin(0.5 * (pi_r / (md + 1)));
    xb *= xb;
    fb = 0.0;
    for (int it = 0; std::abs(xa - xb) > tol && it < 10000; ++it) {
      const double slope = (fb - fa) / (xb - xa);
      const double xc = xb - (pm(m, xb) - y) / slope;
      xa = xb;
      fa = fb;
      xb = xc;
      fb = pm(m, xb);
    }
  }
  if (y > 5.0) {
    xa = 0.0;
    fa = 0.0;
    xb = -4.0 / ((md + 1.0) * (md + 1.0));
    fb = std::log(pm(m, xb));
    for (int it = 0; std::abs(xa - xb) > tol && it < 10000; ++it) {
      const double slope = (fb - fa) / (xb - xa);
      const double xc = xb - (std::log(pm(m, xb)) - std::log(y)) / slope;
      xa = xb;
      fa = fb;
      xb = xc;
      fb = std::log(pm(m, xb));
    }
  }
  return xb;
}

inline double x_star(long m) {
  if (m == 0) return 1.0;
  if (m == 1) return 0.3;
  const double xp = pm_prime_first_root(m);
  return pm_inverse(m, -pm(m, xp));
}

inline double upper_x_star(long m) {
  if (m == 0) return 1.0;
  if (m == 1) return 0.6;
  if (m == 2) return ((1.0 - x_star(m) + 0.6) / 2.0);
  return pm_next_to_last_root(m);
}

// Smallest degree m (>= 1, capped at mMax) whose X*/UpperX* ratio maps
// lambdaStar below X*, and the bound that maps lambdaStar onto X* exactly.
inline void star_degree_and_bound(double shift, double lambda_bound, double lambda_star,
                                  long m_max, long* degree, double* bound) {
  const double ls = lambda_star + shift;
  double ratio = x_star(1) / upper_x_star(1);
  if (ratio * (lambda_bound + shift) < ls) {
    *degree = 1;
    *bound = (ls / ratio) - shift;
    return;
  }
  long m = 2;
  for (; m < m_max; ++m) {
    ratio = x_star(m) / upper_x_star(m);
    if (ratio * (lambda_bound + shift) < ls) break;
  }
  if (m >= m_max) {
    *degree = m_max;
    *bound = lambda_bound;
    return;
  }
  *degree = m;
  ratio = x_star(m) / upper_x_star(m);
  *bound = (ls / ratio) - shift;
}

// Filter coefficients of CPO-LM:204-207 for degree m, bound lmax and shift.
struct Coefs {
  double rhoB, g1, g2, upper;
};
inline Coefs coefficients(long m, double lmax, double shift) {
  Coefs c;
  c.upper = upper_x_star(m);
  c.rhoB = lmax / c.upper + shift / c.upper;
  c.g1 = -2.0 / (c.rhoB - 2.0 * shift);
  c.g2 = 2.0 - (4.0 * shift) / c.rhoB;
  return c;
}

// P_m((lambda + shift)/rhoB)^rep, the scalar the filter applies to an eigenvalue.
inline double evaluate(long m, long rep, double lmax, double shift, double lambda) {
  const Coefs c = coefficients(m, lmax, shift);
  return std::pow(pm(m, (lambda + shift) / c.rhoB), int(rep));
}

}  // namespace cpoly
}  // namespace rcg
