// rc_oracle_capi.cpp -- extern "C" surface of the CPU oracle (liborc.so).
//
// TEST INFRASTRUCTURE ONLY: loaded with ctypes by tests/, the smoke check and
// the CPU-baseline leg of bench.py.  Every function follows rc_oracle.hpp,
// which cites the reference file:line it restates.
#include <cstring>
#include <exception>
#include <string>

#include "rc_oracle.hpp"

using orc::cplx;

namespace {
thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return 2;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

template <class T>
orc::Block<T> cols_to_block(const T* X, long n, long b) {
  orc::Block<T> V(b, std::vector<T>(n));
  for (long k = 0; k < b; ++k) std::memcpy(V[k].data(), X + k * n, sizeof(T) * n);
  return V;
}
template <class T>
void block_to_cols(const orc::Block<T>& V, T* X, long n) {
  for (size_t k = 0; k < V.size(); ++k) std::memcpy(X + k * n, V[k].data(), sizeof(T) * n);
}
}  // namespace

extern "C" {

struct orc_params {
  long k;
  long inc;
  long buf;
  double tol;
  unsigned long long seed;
  int stop;             // 0 DEFAULT, 1 COMBINATION, 2 EIGENVALUE_ONLY, 3 RESIDUAL_ONLY
  int interval;         // 1: getMinIntervalEigenSystem(lambda_max, ...)
  long max_inner;       // maxInnerLoopCount
  long deg_max;         // minIntervalPolyDegreeMax
  long rep;             // filterRepetitionCount
  double lambda_max;    // interval bound
  long lanczos_itmax;   // setMaxSpectralBoundsIter
  double maxmin_tol;    // setMaxMinTolerance
  int fixed_iteration;  // setFixedIterationCount
  int pad;
  long out_cap;         // capacity (in eigenpairs) of evals/resid/evecs
};

struct orc_result {
  long found;
  long lanczos_iters;
  double min_est;
  double max_est;
  double guard;
  long count_ortho;
  long count_apply;
  long count_eig;
  double final_maxres;
  double final_maxdiff;
  double final_maxgap;
  long trace_len;
  unsigned long long rng_counter;
  double t_lanczos, t_filter, t_ortho, t_deflate, t_rr;  // wall seconds per phase
  long n_ortho, n_deflate, n_rr, degree_sum;
};

const char* orc_last_error() { return g_err.c_str(); }

double orc_unit_draw(unsigned long long seed, unsigned long long idx) {
  return orc::unit_draw(seed, idx);
}
void orc_fill_sym(unsigned long long seed, unsigned long long counter, double* out, long n) {
  for (long i = 0; i < n; ++i) out[i] = orc::sym_draw(seed, counter + i);
}

double orc_cpoly_pm(long m, double x) {
  orc::Cpoly c;
  c.polyDegree = m;
  return c.evaluatePm(x);
}
double orc_cpoly_pm_prime(long m, double x) {
  orc::Cpoly c;
  c.polyDegree = m;
  return c.evaluatePmPrime(x);
}
double orc_cpoly_xstar(long m) {
  orc::Cpoly c;
  c.polyDegree = m;
  return c.getXStar();
}
double orc_cpoly_upper_xstar(long m) {
  orc::Cpoly c;
  c.polyDegree = m;
  return c.getUpperXStar();
}
double orc_cpoly_inverse_pm(long m, double y) {
  orc::Cpoly c;
  c.polyDegree = m;
  return c.evaluateInversePm(y);
}
double orc_cpoly_pm_prime_first_root(long m) {
  orc::Cpoly c;
  c.polyDegree = m;
  return c.getPmPrimeFirstRoot();
}
void orc_cpoly_star_degree(double shift, double lambdaBound, double lambdaStar, long degMax,
                           long* deg, double* bound) {
  orc::Cpoly c;
  c.getStarDegreeAndSpectralRadius(shift, lambdaBound, lambdaStar, degMax, *deg, *bound);
}
double orc_cpoly_eval_polynomial(long m, long rep, double lmax, double shift, double lambda) {
  orc::Cpoly c;
  c.set(m, rep, lmax, shift);
  return c.evaluatePolynomial(lambda);
}

double orc_tri_eig(const double* D, const double* E, long n, long idx, double abstol) {
  std::vector<double> d(D, D + n), e(E, E + (n > 0 ? n - 1 : 0));
  return orc::tri_eig(d, e, idx, abstol);
}

void orc_jacobi_f64(const double* H, long n, double* evals, double* evecs) {
  orc::Jacobi J;
  J.eigensystem(H, n, evals, evecs);
}
void orc_jacobi_c128(const double* H, long n, double* evals, double* evecs) {
  orc::HermitianJacobi::eigensystem(reinterpret_cast<const cplx*>(H), n, evals,
                                    reinterpret_cast<cplx*>(evecs));
}

// Modified Gram-Schmidt of a column-major n x b block (RayleighChebyshevLM.h:1782-1837).
void orc_mgs_f64(double* X, long n, long b) {
  auto V = cols_to_block(X, n, b);
  orc::RCLM<double, orc::CsrOp<double>>::orthogonalize(V);
  block_to_cols(V, X, n);
}

// Block filter on a column-major n x b block, CSR operator.
int orc_filter_csr_f64(long n, const int* rowptr, const int* col, const double* val, double* X,
                       long b, long m, long rep, double lambdaMax, double shift) {
  return guarded([&] {
    orc::CsrOp<double> op{n, rowptr, col, val};
    orc::Cpoly c;
    c.set(m, rep, lambdaMax, shift);
    auto V = cols_to_block(X, n, b);
    orc::filter_apply<double>(op, V, m, rep, lambdaMax, shift, c.UpperXStar);
    block_to_cols(V, X, n);
  });
}
int orc_filter_csr_c128(long n, const int* rowptr, const int* col, const double* val, double* X,
                        long b, long m, long rep, double lambdaMax, double shift) {
  return guarded([&] {
    orc::CsrOp<cplx> op{n, rowptr, col, reinterpret_cast<const cplx*>(val)};
    orc::Cpoly c;
    c.set(m, rep, lambdaMax, shift);
    auto V = cols_to_block(reinterpret_cast<cplx*>(X), n, b);
    orc::filter_apply<cplx>(op, V, m, rep, lambdaMax, shift, c.UpperXStar);
    block_to_cols(V, reinterpret_cast<cplx*>(X), n);
  });
}
int orc_filter_dense_f64(long n, const double* A, double* X, long b, long m, long rep,
                         double lambdaMax, double shift) {
  return guarded([&] {
    orc::DenseOp<double> op{n, A};
    orc::Cpoly c;
    c.set(m, rep, lambdaMax, shift);
    auto V = cols_to_block(X, n, b);
    orc::filter_apply<double>(op, V, m, rep, lambdaMax, shift, c.UpperXStar);
    block_to_cols(V, X, n);
  });
}

int orc_lanczos_csr_f64(long n, const int* rowptr, const int* col, const double* val,
                        unsigned long long seed, double tol, int maxFlag, int minFlag, long itmax,
                        double* mn, double* mx, long* iters) {
  return guarded([&] {
    orc::CsrOp<double> op{n, rowptr, col, val};
    orc::RandomStream rnd{seed, 0};
    std::vector<double> v(n);
    auto r = orc::lanczos_min_max<double>(op, v, rnd, tol, maxFlag != 0, minFlag != 0, itmax);
    *mn = r.minEig;
    *mx = r.maxEig;
    *iters = r.iters;
  });
}

}  // extern "C"

namespace {
template <class T, class Op>
int solve_impl(const Op& op, long n, const orc_params* p, orc_result* r, double* evals,
               T* evecs, double* resid, double* trace, long trace_cap) {
  return guarded([&] {
    orc::RCLM<T, Op> s;
    s.stopCondition = orc::StopCondition(p->stop == 0 ? 1 : p->stop);
    if (p->max_inner > 0) s.maxInnerLoopCount = p->max_inner;
    if (p->deg_max > 0) s.minIntervalPolyDegreeMax = p->deg_max;
    if (p->rep > 0) s.filterRepetitionCount = p->rep;
    if (p->lanczos_itmax > 0) s.lanczosIterMax = p->lanczos_itmax;
    if (p->maxmin_tol > 0) s.maxMinTol = p->maxmin_tol;
    s.fixedIterationCount = p->fixed_iteration != 0;
    orc::RandomStream rnd{p->seed, 0};
    std::vector<T> vStart(n);
    std::vector<double> ev;
    orc::Block<T> V;
    long found;
    if (p->interval)
      found = s.getMinIntervalEigenSystem(p->lambda_max, p->tol, p->inc, p->buf, p->k, vStart, op,
                                          rnd, ev, V);
    else
      found = s.getMinEigenSystem(p->k, p->tol, p->inc, p->buf, vStart, op, rnd, ev, V);
    r->found = found;
    r->lanczos_iters = s.lanczosIters;
    r->min_est = s.minEigValueEst;
    r->max_est = s.maxEigValueEst;
    r->guard = s.guardValue;
    r->count_ortho = s.countData["ortho"];
    r->count_apply = s.countData["OpApply"];
    r->count_eig = s.countData["eigenvalue"];
    r->final_maxres = s.finalData["maxResidual"];
    r->final_maxdiff = s.finalData["maxEigValueDifference"];
    r->final_maxgap = s.finalData["maxRelEigValueGap"];
    r->trace_len = (long)s.trace.size();
    r->rng_counter = rnd.counter;
    r->t_lanczos = s.tLanczos;
    r->t_filter = s.tFilter;
    r->t_ortho = s.tOrtho;
    r->t_deflate = s.tDeflate;
    r->t_rr = s.tRR;
    r->n_ortho = s.nOrtho;
    r->n_deflate = s.nDeflate;
    r->n_rr = s.nRR;
    r->degree_sum = s.degreeSum;
    const long cap = p->out_cap;
    for (long i = 0; i < (long)ev.size() && i < cap; ++i) evals[i] = ev[i];
    if (resid)
      for (long i = 0; i < (long)s.eigVecResiduals.size() && i < cap; ++i)
        resid[i] = s.eigVecResiduals[i];
    if (evecs)
      for (long i = 0; i < (long)V.size() && i < cap; ++i)
        std::memcpy(evecs + i * n, V[i].data(), sizeof(T) * n);
    if (trace)
      for (long t = 0; t < (long)s.trace.size() && t < trace_cap; ++t) {
        const auto& row = s.trace[t];
        double* o = trace + 10 * t;
        o[0] = row.pass; o[1] = row.iter; o[2] = row.degree; o[3] = row.starBound;
        o[4] = row.shift; o[5] = row.lambdaStar; o[6] = row.maxResidual; o[7] = row.maxEigDiff;
        o[8] = row.eigDiffRatio; o[9] = row.maxGap;
      }
  });
}
}  // namespace

extern "C" {

// Lowest-eigenpair solve, RayleighChebyshevLM::getMinEigenSystem (or the
// interval variant).  evals/resid sized k, evecs (optional) n*k column-major.
int orc_solve_csr_f64(long n, const int* rowptr, const int* col, const double* val,
                      const orc_params* p, orc_result* r, double* evals, double* evecs,
                      double* resid, double* trace, long trace_cap) {
  orc::CsrOp<double> op{n, rowptr, col, val};
  return solve_impl<double>(op, n, p, r, evals, evecs, resid, trace, trace_cap);
}
int orc_solve_csr_c128(long n, const int* rowptr, const int* col, const double* val,
                       const orc_params* p, orc_result* r, double* evals, double* evecs,
                       double* resid, double* trace, long trace_cap) {
  orc::CsrOp<cplx> op{n, rowptr, col, reinterpret_cast<const cplx*>(val)};
  return solve_impl<cplx>(op, n, p, r, evals, reinterpret_cast<cplx*>(evecs), resid, trace,
                          trace_cap);
}
int orc_solve_dense_f64(long n, const double* A, const orc_params* p, orc_result* r,
                        double* evals, double* evecs, double* resid, double* trace,
                        long trace_cap) {
  orc::DenseOp<double> op{n, A};
  return solve_impl<double>(op, n, p, r, evals, evecs, resid, trace, trace_cap);
}

long orc_sizeof_params() { return (long)sizeof(orc_params); }
long orc_sizeof_result() { return (long)sizeof(orc_result); }

}  // extern "C"
