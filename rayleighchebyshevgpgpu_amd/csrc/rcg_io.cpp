// rcg_io.cpp -- operator ingest on the host: Matrix Market files and 64-bit
// CSR index arrays (the reference's RAY_INT width, RayleighChebyshevLMCuda.h
// :409ff), turned into the int32 CSR the device kernels read.
//
// Matrix Market coordinate files: real, integer, complex or pattern values;
// general, symmetric or hermitian storage (the stored triangle is mirrored,
// conjugated for hermitian).  Rows are sorted by column and duplicate entries
// summed, as assemblers expect.
#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/rcg_c_api.h"
#include "rcg_internal.hpp"

namespace {

struct MmEntry {
  int64_t r, c;
  double re, im;
};

struct MmMatrix {
  int64_t n = 0;
  bool complex = false;
  std::vector<int32_t> rowptr, col;
  std::vector<double> val;  // re (and im interleaved when complex)
};

std::string lower(std::string s) {
  for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

void fail(const std::string& m) { throw rcg::Error(RCG_ERR_INVALID, m); }

MmMatrix read_mm(const char* path) {
  std::ifstream in(path);
  if (!in) fail(std::string("Matrix Market: cannot open ") + path);
  std::string line;
  if (!std::getline(in, line)) fail("empty file");
  std::istringstream hs(line);
  std::string banner, object, format, field, symmetry;
  hs >> banner >> object >> format >> field >> symmetry;
  if (lower(banner) != "%%matrixmarket" || lower(object) != "matrix")
    fail("not a MatrixMarket matrix file");
  if (lower(format) != "coordinate") fail("only the coordinate format is read");
  field = lower(field);
  symmetry = lower(symmetry);
  const bool pattern = field == "pattern";
  const bool cplx = field == "complex";
  if (!pattern && !cplx && field != "real" && field != "integer" && field != "double")
    fail("unsupported field " + field);
  const bool sym = symmetry == "symmetric", herm = symmetry == "hermitian";
  if (!sym && !herm && symmetry != "general") fail("unsupported symmetry " + symmetry);
  while (std::getline(in, line))
    if (!line.empty() && line[0] != '%') break;
  std::istringstream ss(line);
  int64_t m = 0, n = 0, nz = 0;
  if (!(ss >> m >> n >> nz)) fail("bad size line");
  if (m != n || n <= 0) fail("the operator must be square");
  if (n > INT32_MAX) fail("dimension exceeds the int32 index range");
  std::vector<MmEntry> e;
  e.reserve((size_t)nz * ((sym || herm) ? 2 : 1));
  for (int64_t k = 0; k < nz; ++k) {
    if (!std::getline(in, line)) fail("fewer entries than declared");
    if (line.empty() || line[0] == '%') {
      --k;
      continue;
    }
    std::istringstream es(line);
    MmEntry t{0, 0, 1.0, 0.0};
    if (!(es >> t.r >> t.c)) fail("bad entry line");
    if (!pattern && !(es >> t.re)) fail("missing value");
    if (cplx && !(es >> t.im)) fail("missing imaginary part");
    if (t.r < 1 || t.r > n || t.c < 1 || t.c > n) fail("entry index out of range");
    --t.r;
    --t.c;
    e.push_back(t);
    if ((sym || herm) && t.r != t.c) e.push_back({t.c, t.r, t.re, herm ? -t.im : t.im});
  }
  std::sort(e.begin(), e.end(), [](const MmEntry& a, const MmEntry& b) {
    return a.r != b.r ? a.r < b.r : a.c < b.c;
  });
  MmMatrix M;
  M.n = n;
  M.complex = cplx;
  M.rowptr.assign(n + 1, 0);
  for (size_t k = 0; k < e.size();) {
    size_t j = k;
    double re = 0.0, im = 0.0;
    while (j < e.size() && e[j].r == e[k].r && e[j].c == e[k].c) {
      re += e[j].re;
      im += e[j].im;
      ++j;
    }
    M.col.push_back((int32_t)e[k].c);
    M.val.push_back(re);
    if (cplx) M.val.push_back(im);
    M.rowptr[e[k].r + 1]++;
    k = j;
  }
  for (int64_t i = 0; i < n; ++i) M.rowptr[i + 1] += M.rowptr[i];
  if ((int64_t)M.col.size() > INT32_MAX) fail("nonzeros exceed the int32 index range");
  return M;
}

template <class F>
int guarded_io(F&& f) {
  try {
    f();
    return RCG_OK;
  } catch (const rcg::Error& e) {
    rcg::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    rcg::set_last_error(e.what());
    return RCG_ERR_INVALID;
  }
}

}  // namespace

extern "C" {

int rcg_matrix_market_info(const char* path, int64_t* n, int64_t* nnz, int* dtype) {
  return guarded_io([&] {
    if (!path || !n || !nnz || !dtype) fail("NULL argument");
    const MmMatrix M = read_mm(path);
    *n = M.n;
    *nnz = (int64_t)M.col.size();
    *dtype = M.complex ? RCG_C128 : RCG_F64;
  });
}

int rcg_matrix_market_csr(const char* path, int32_t* rowptr, int32_t* col, void* val) {
  return guarded_io([&] {
    if (!path || !rowptr || !col || !val) fail("NULL argument");
    const MmMatrix M = read_mm(path);
    std::memcpy(rowptr, M.rowptr.data(), sizeof(int32_t) * M.rowptr.size());
    std::memcpy(col, M.col.data(), sizeof(int32_t) * M.col.size());
    std::memcpy(val, M.val.data(), sizeof(double) * M.val.size());
  });
}

int rcg_operator_from_matrix_market(rcg_context ctx, const char* path, rcg_operator* out) {
  MmMatrix M;
  int rc = guarded_io([&] {
    if (!ctx || !path || !out) fail("NULL argument");
    M = read_mm(path);
  });
  if (rc != RCG_OK) return rc;
  return rcg_operator_csr_create(ctx, M.n, (int64_t)M.col.size(), M.rowptr.data(), M.col.data(),
                                 M.val.data(), M.complex ? RCG_C128 : RCG_F64, RCG_MEM_HOST, out);
}

int rcg_operator_csr64_create(rcg_context ctx, int64_t n, int64_t nnz, const int64_t* rowptr,
                              const int64_t* col, const void* val, int val_dtype,
                              rcg_operator* out) {
  std::vector<int32_t> rp, ci;
  int rc = guarded_io([&] {
    if (!ctx || !rowptr || !col || !val || !out) fail("NULL argument");
    if (n <= 0 || nnz < 0) fail("bad operator size");
    if (n > INT32_MAX || nnz > INT32_MAX)
      throw rcg::Error(RCG_ERR_UNSUPPORTED, "sizes exceed the int32 index range of the kernels");
    rp.resize(n + 1);
    ci.resize(nnz);
    for (int64_t i = 0; i <= n; ++i) {
      if (rowptr[i] < 0 || rowptr[i] > nnz) fail("row pointer out of range");
      rp[i] = (int32_t)rowptr[i];
    }
    for (int64_t p = 0; p < nnz; ++p) {
      if (col[p] < 0 || col[p] >= n) fail("column index out of range");
      ci[p] = (int32_t)col[p];
    }
  });
  if (rc != RCG_OK) return rc;
  return rcg_operator_csr_create(ctx, n, nnz, rp.data(), ci.data(), val, val_dtype, RCG_MEM_HOST,
                                 out);
}

}  // extern "C"
