// ref_harness.cpp -- caller-side harness for the compiled reference.
//
// TEST INFRASTRUCTURE ONLY.  The reference ships headers and no main; this
// file is the "user code" its documentation asks for (RayleighChebyshevLM.h:
// 36-124): a CSR operator with apply(std::vector<Vtype>&) and apply(Vtype&),
// and a start-vector operator with randomize(Vtype&).  It includes the
// reference headers where they lie (-I to the reference checkout) and is
// built by oracle/Makefile into oracle/_ref/ only; it never ships to the GPU.
// The vector type is the reference's own AvectorClassLMCuda<double>.
//
// usage: rc_ref <csr.bin> <out.json> <k> <tol> <inc> <buf> <seed> <stop 0-3>
//               <mode lowest|interval> <lambdaMax> <maxInner>
// csr.bin: int64 n, int64 nnz, int32 rowptr[n+1], int32 col[nnz], float64 val[nnz]
#include <complex>
#include <cstdint>
#include <cstdio>
#include <type_traits>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "AvectorClassLMCuda.h"
#include "RayleighChebyshevLM.h"

using Vec = AvectorClassLMCuda<double>;

namespace {
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ULL;
uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double sym_draw(uint64_t seed, uint64_t idx) {
  const uint64_t key = mix64(seed + kGamma);
  const uint64_t z = mix64(key + (idx + 1) * kGamma);
  return -1.0 + 2.0 * (double(z >> 11) * 0x1.0p-53);
}
}  // namespace

struct CounterRandomize {
  uint64_t seed = 0;
  uint64_t counter = 0;
  void randomize(Vec& v) {
    for (size_t i = 0; i < v.vData.size(); ++i) v.vData[i] = sym_draw(seed, counter + i);
    counter += v.vData.size();
  }
};

struct CsrOperator {
  long n = 0;
  std::vector<int> rowptr, col;
  std::vector<double> val;
  CsrOperator() = default;
  CsrOperator(const CsrOperator&) = default;

  void apply(Vec& v) {
    std::vector<double> y(n);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
      double s = 0.0;
      for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) s += val[p] * v.vData[col[p]];
      y[i] = s;
    }
    v.vData.swap(y);
  }
  void apply(std::vector<Vec>& V) {
    const long b = (long)V.size();
    std::vector<std::vector<double>> Y(b, std::vector<double>(n));
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
      for (long k = 0; k < b; ++k) {
        double s = 0.0;
        const double* x = V[k].vData.data();
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) s += val[p] * x[col[p]];
        Y[k][i] = s;
      }
    }
    for (long k = 0; k < b; ++k) V[k].vData.swap(Y[k]);
  }
};

int main(int argc, char** argv) {
  if (argc < 12) {
    std::fprintf(stderr, "usage: rc_ref csr.bin out.json k tol inc buf seed stop mode lambdaMax maxInner\n");
    return 2;
  }
  CsrOperator op;
  {
    std::ifstream f(argv[1], std::ios::binary);
    int64_t n = 0, nnz = 0;
    f.read(reinterpret_cast<char*>(&n), 8);
    f.read(reinterpret_cast<char*>(&nnz), 8);
    op.n = n;
    op.rowptr.resize(n + 1);
    op.col.resize(nnz);
    op.val.resize(nnz);
    f.read(reinterpret_cast<char*>(op.rowptr.data()), 4 * (n + 1));
    f.read(reinterpret_cast<char*>(op.col.data()), 4 * nnz);
    f.read(reinterpret_cast<char*>(op.val.data()), 8 * nnz);
    if (!f) {
      std::fprintf(stderr, "rc_ref: short read of %s\n", argv[1]);
      return 2;
    }
  }
  const long k = std::atol(argv[3]);
  const double tol = std::atof(argv[4]);
  const long inc = std::atol(argv[5]);
  const long buf = std::atol(argv[6]);
  const uint64_t seed = std::strtoull(argv[7], nullptr, 10);
  const int stop = std::atoi(argv[8]);
  const std::string mode = argv[9];
  const double lambdaMax = std::atof(argv[10]);
  const long maxInner = std::atol(argv[11]);

  RayleighChebyshevLM<Vec, CsrOperator, CounterRandomize, double> rc;
  rc.setStopCondition(RC_Types::StopCondition(stop));
  if (maxInner > 0) rc.setMaxInnerLoopCount(maxInner);
  rc.setVerboseFlag(true);

  CounterRandomize rnd;
  rnd.seed = seed;
  Vec vStart(op.n);
  std::vector<double> evals;
  std::vector<Vec> evecs;

  std::ostringstream captured;
  std::streambuf* old = std::cout.rdbuf(captured.rdbuf());
  long found = 0;
  std::string err;
  try {
    if (mode == "interval")
      found = rc.getMinIntervalEigenSystem(lambdaMax, tol, inc, buf, k, vStart, op, rnd, evals, evecs);
    else
      found = rc.getMinEigenSystem(k, tol, inc, buf, vStart, op, rnd, evals, evecs);
  } catch (const std::exception& e) {
    err = e.what();
  }
  std::cout.rdbuf(old);

  std::ofstream o(argv[2]);
  o.precision(17);
  o << "{\n  \"found\": " << found << ",\n  \"error\": \"" << (err.empty() ? "" : "exception") << "\"";
  o << ",\n  \"eigenvalues\": [";
  for (size_t i = 0; i < evals.size(); ++i) o << (i ? ", " : "") << evals[i];
  o << "],\n  \"residuals\": [";
  auto res = rc.getEigVectorResiduals();
  for (size_t i = 0; i < res.size(); ++i) o << (i ? ", " : "") << res[i];
  o << "],\n  \"lanczos_iters\": " << rc.getSpectralBoundsIterCount();
  o << ",\n  \"min_est\": " << rc.getMinEigValueEst() << ",\n  \"max_est\": " << rc.getMaxEigValueEst();
  o << ",\n  \"guard\": " << rc.getGuardEigenvalue();
  auto cd = rc.getCountData();
  o << ",\n  \"count\": {\"ortho\": " << cd["ortho"] << ", \"OpApply\": " << cd["OpApply"]
    << ", \"eigenvalue\": " << cd["eigenvalue"] << "}";
  auto fd = rc.getFinalData();
  o << ",\n  \"final\": {\"maxResidual\": " << fd["maxResidual"]
    << ", \"maxEigValueDifference\": " << fd["maxEigValueDifference"]
    << ", \"maxRelEigValueGap\": " << fd["maxRelEigValueGap"] << "}";
  o << ",\n  \"rng_counter\": " << rnd.counter;
  // Per-iteration verbose lines: "<iter> : Degree <m>  Residual Max: <r>  Eig Diff Max: <d> ..."
  o << ",\n  \"trace\": [";
  std::istringstream in(captured.str());
  std::string line;
  bool first = true;
  while (std::getline(in, line)) {
    const auto pd = line.find(": Degree ");
    if (pd == std::string::npos) continue;
    long it = 0, deg = 0;
    double rmax = 0, dmax = 0, ratio = 0, gap = 0;
    if (std::sscanf(line.c_str(), "%ld : Degree %ld Residual Max: %lf Eig Diff Max: %lf Eig Conv Factor: %lf Max Gap %lf",
                    &it, &deg, &rmax, &dmax, &ratio, &gap) == 6) {
      o << (first ? "" : ", ") << "[" << it << ", " << deg << ", " << rmax << ", " << dmax << ", "
        << ratio << ", " << gap << "]";
      first = false;
    }
  }
  o << "]\n}\n";
  return err.empty() ? 0 : 1;
}
