// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (built and
// run by tests/test_sanitizers.py, no GPU needed): the Matrix Market parser
// of rcg_io.cpp on well-formed and malformed files, and the CPU oracle
// (oracle/rc_oracle_capi.cpp, compiled into this program) solving a 2-D
// Dirichlet Laplacian, real and as a complex (Hermitian) operator.
//
//   usage: ingest_oracle_asan <dir> <file.mtx>...
// prints one line per file ("name status n nnz [csr-status]") and one line
// per oracle solve ("oracle <kind> found <f> max|dlambda| <d>"); exits
// non-zero on a wrong result, and a sanitizer report aborts the run.
#include <cmath>
#include <complex>
#include <cstdio>
#include <string>
#include <vector>

#include "rcg_c_api.h"

#include "../../oracle/rc_oracle_capi.cpp"

namespace {

// 2-D Dirichlet 5-point Laplacian on nx x nx, columns sorted
void laplacian(long nx, std::vector<int>& rp, std::vector<int>& ci, std::vector<double>& v) {
  rp.assign(1, 0);
  ci.clear();
  v.clear();
  for (long y = 0; y < nx; ++y)
    for (long x = 0; x < nx; ++x) {
      const long r = x + nx * y;
      const long nb[5] = {y > 0 ? r - nx : -1, x > 0 ? r - 1 : -1, r, x + 1 < nx ? r + 1 : -1,
                          y + 1 < nx ? r + nx : -1};
      for (long c : nb)
        if (c >= 0) {
          ci.push_back((int)c);
          v.push_back(c == r ? 4.0 : -1.0);
        }
      rp.push_back((int)ci.size());
    }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  int bad = 0;
  for (int i = 2; i < argc; ++i) {
    const std::string path = dir + "/" + argv[i];
    int64_t n = 0, nnz = 0;
    int dt = 0;
    const int st = rcg_matrix_market_info(path.c_str(), &n, &nnz, &dt);
    std::printf("%s %d %lld %lld", argv[i], st, (long long)n, (long long)nnz);
    if (st == 0) {
      std::vector<int32_t> rp((size_t)n + 1), col((size_t)(nnz > 0 ? nnz : 1));
      std::vector<double> val((size_t)(dt == RCG_C128 ? 2 : 1) * (size_t)(nnz > 0 ? nnz : 1));
      const int st2 = rcg_matrix_market_csr(path.c_str(), rp.data(), col.data(), val.data());
      std::printf(" %d", st2);
      if (st2 == 0 && (rp[0] != 0 || rp[(size_t)n] != nnz)) bad = 1;
    }
    std::printf("\n");
  }

  const long nx = 16, n = nx * nx, k = 6;
  std::vector<int> rp, ci;
  std::vector<double> v;
  laplacian(nx, rp, ci, v);
  std::vector<double> exact;
  const double pi = std::acos(-1.0);
  for (long a = 1; a <= nx; ++a)
    for (long b = 1; b <= nx; ++b)
      exact.push_back(4.0 - 2.0 * std::cos(pi * a / (nx + 1)) - 2.0 * std::cos(pi * b / (nx + 1)));
  std::sort(exact.begin(), exact.end());
  orc_params p{};
  p.k = k;
  p.inc = k;
  p.buf = 4;
  p.tol = 1e-10;
  p.stop = 3;
  p.out_cap = k + 4;
  std::vector<double> ev(k + 4), res(k + 4), trace(10 * 4096);
  for (int cx = 0; cx < 2; ++cx) {
    orc_result r{};
    int st;
    if (cx) {  // the same operator with complex values
      std::vector<double> vc(2 * v.size(), 0.0);
      for (size_t q = 0; q < v.size(); ++q) vc[2 * q] = v[q];
      st = orc_solve_csr_c128(n, rp.data(), ci.data(), vc.data(), &p, &r, ev.data(), nullptr,
                              res.data(), trace.data(), 4096);
    } else {
      st = orc_solve_csr_f64(n, rp.data(), ci.data(), v.data(), &p, &r, ev.data(), nullptr,
                             res.data(), trace.data(), 4096);
    }
    double d = 0.0;
    for (long j = 0; j < r.found && j < k; ++j) d = std::max(d, std::fabs(ev[j] - exact[j]));
    std::printf("oracle %s status %d found %ld max|dlambda| %.3e\n", cx ? "c128" : "f64", st,
                r.found, d);
    if (st != 0 || r.found < k || d > 1e-9) bad = 1;
  }
  return bad;
}
