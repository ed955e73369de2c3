// RayleighChebyshevLM.hpp -- C++ host mirror of the reference's template API
// over the C ABI of librcg.so (include/rcg_c_api.h).
//
// A caller of the reference writes
//
//   RayleighChebyshevLMCuda<Atype, Vtype, Otype, RandOp, double> rc;
//   rc.setStopCondition("RESIDUAL_ONLY");
//   long found = rc.getMinEigenSystem(k, tol, inc, buf, vStart, oP, randOp,
//                                     eigValues, eigVectors);
//
// (RayleighChebyshevLMCuda.h:220-2472, RayleighChebyshevLM.h:213-2246).  The
// same code compiles against this header with the rcg types below standing in
// for the template arguments: AmatrixClass (AmatrixClassLMCuda.h: a device
// N x n block with a column-major host mirror), AvectorClass
// (AvectorClassLMCuda.h: a host vector, here only the dimension template for
// vStart), CsrOperator (the user Otype with apply(in, out, alpha, beta)) and
// ArandomizeOpClass (ArandomizeOpClassLMCuda.h, with a seed).  The whole solve
// runs on the GPU inside librcg.so; API misuse throws std::runtime_error and a
// non-Hermitian operator std::invalid_argument, as in the reference.
#pragma once

#include <algorithm>
#include <cctype>
#include <cmath>
#include <complex>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../rcg_c_api.h"

namespace rcg {

inline void check(int status) {
  if (status == RCG_OK) return;
  char buf[2048];
  rcg_get_last_error(buf, sizeof(buf));
  if (status == RCG_ERR_NOT_HERMITIAN) throw std::invalid_argument(buf);
  throw std::runtime_error(std::string("rcg: ") + buf);
}

// Device + stream binding (one per GPU and stream).
class Context {
 public:
  explicit Context(int device = 0, void* hip_stream = nullptr) {
    if (rcg_sizeof_solve_params() != (int64_t)sizeof(rcg_solve_params) ||
        rcg_sizeof_solve_result() != (int64_t)sizeof(rcg_solve_result))
      throw std::runtime_error("rcg: librcg.so and rcg_c_api.h disagree on the solve structs");
    check(rcg_context_create(device, hip_stream, &h_));
  }
  ~Context() { rcg_context_destroy(h_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  void synchronize() { check(rcg_context_synchronize(h_)); }
  rcg_context handle() const { return h_; }

 private:
  rcg_context h_ = nullptr;
};

// AvectorClassLMCuda role (AvectorClassLMCuda.h:43-143): a host vector.  The
// solver only reads its dimension (vStart is the template for eigenvectors).
template <typename T = double>
class AvectorClass {
 public:
  AvectorClass() = default;
  explicit AvectorClass(long dimension) : vData(dimension) {}
  void initialize(const AvectorClass& W) { vData = W.vData; }
  T dot(const AvectorClass& W) const {
    T s = T(0);
    for (size_t k = 0; k < vData.size(); ++k) s += vData[k] * conj_(W.vData[k]);
    return s;
  }
  void operator*=(T a) { for (auto& x : vData) x *= a; }
  void operator+=(const AvectorClass& W) { for (size_t k = 0; k < vData.size(); ++k) vData[k] += W.vData[k]; }
  void operator-=(const AvectorClass& W) { for (size_t k = 0; k < vData.size(); ++k) vData[k] -= W.vData[k]; }
  T operator[](long k) const { return vData[k]; }
  T norm2() const { return std::sqrt(std::abs(dot(*this))); }
  size_t getDimension() const { return vData.size(); }
  std::vector<T> vData;

 private:
  static double conj_(double x) { return x; }
  static std::complex<double> conj_(std::complex<double> x) { return std::conj(x); }
};

// AmatrixClassLMCuda role (AmatrixClassLMCuda.h:12-359): an m x n device block
// with a column-major host mirror; host_to_device / device_to_host copy
// between them.  The device layout is row-interleaved (see rcg_c_api.h).
template <typename T = double>
class AmatrixClass {
 public:
  AmatrixClass(Context& ctx, long m, long n) : ctx_(&ctx) { resize(m, n); }
  ~AmatrixClass() { release(); }
  AmatrixClass(const AmatrixClass&) = delete;
  AmatrixClass& operator=(const AmatrixClass&) = delete;

  void resize(long m, long n) {
    if (m == m_ && n == n_ && h_) return;
    release();
    m_ = m;
    n_ = n;
    mData.assign((size_t)m * (size_t)n, T(0));
    check(rcg_block_create(ctx_->handle(), m, n, dtype(), &h_));
  }
  void resize(long n) { resize(m_, n); }
  void host_to_device() { check(rcg_block_upload(h_, mData.data(), m_, n_)); }
  void device_to_host() { check(rcg_block_download(h_, mData.data(), m_, n_)); }
  T& operator()(long i, long j) { return mData[i + j * m_]; }
  const T& operator()(long i, long j) const { return mData[i + j * m_]; }
  long getRowSize() const { return m_; }
  long getColSize() const { return n_; }
  size_t getDimension() const { return (size_t)m_ * (size_t)n_; }
  T* getDataPointer() { return mData.data(); }
  void* getDataPointer_d() const {
    void* p = nullptr;
    check(rcg_block_info(h_, nullptr, nullptr, nullptr, nullptr, &p));
    return p;
  }
  rcg_block handle() const { return h_; }
  long validColumns() const { return valid_; }
  void setValidColumns(long v) { valid_ = v; }

  std::vector<T> mData;

 private:
  static int dtype() {
    if (std::is_same<T, std::complex<double>>::value) return RCG_C128;
    return std::is_same<T, float>::value ? RCG_F32 : RCG_F64;
  }
  void release() {
    if (h_) rcg_block_destroy(h_);
    h_ = nullptr;
  }
  Context* ctx_;
  rcg_block h_ = nullptr;
  long m_ = 0, n_ = 0, valid_ = 0;
};

// The user operator of the reference (Otype with apply(in, out, alpha, beta),
// LanczosCpolyOperatorLMCuda.h:236,245): a CSR matrix on the device.
class CsrOperator {
 public:
  CsrOperator(Context& ctx, long n, const std::vector<int32_t>& rowptr,
              const std::vector<int32_t>& col, const std::vector<double>& val)
      : ctx_(&ctx) {
    if ((long)rowptr.size() != n + 1) throw std::runtime_error("rowptr must have n+1 entries");
    check(rcg_operator_csr_create(ctx.handle(), n, (int64_t)col.size(), rowptr.data(), col.data(),
                                  val.data(), RCG_F64, RCG_MEM_HOST, &h_));
    n_ = n;
  }
  ~CsrOperator() { rcg_operator_destroy(h_); }
  // Complex Hermitian operator (values as std::complex<double>).
  CsrOperator(Context& ctx, long n, const std::vector<int32_t>& rowptr,
              const std::vector<int32_t>& col, const std::vector<std::complex<double>>& val)
      : ctx_(&ctx) {
    if ((long)rowptr.size() != n + 1) throw std::runtime_error("rowptr must have n+1 entries");
    check(rcg_operator_csr_create(ctx.handle(), n, (int64_t)col.size(), rowptr.data(), col.data(),
                                  val.data(), RCG_C128, RCG_MEM_HOST, &h_));
    n_ = n;
  }
  CsrOperator(const CsrOperator&) = delete;
  CsrOperator& operator=(const CsrOperator&) = delete;
  // Locality hint for operators on a structured grid (rows x + nx(y + ny z)).
  void setGrid(long nx, long ny, long nz) { check(rcg_operator_set_grid(h_, nx, ny, nz)); }
  template <class Atype>
  void apply(Atype& in, Atype& out, double alpha, double beta) {
    check(rcg_operator_apply(h_, in.handle(), out.handle(), alpha, beta, in.getColSize()));
  }
  rcg_operator handle() const { return h_; }
  long getDimension() const { return n_; }
  Context& context() const { return *ctx_; }

 private:
  Context* ctx_;
  rcg_operator h_ = nullptr;
  long n_ = 0;
};

// The device block a caller's operator sees inside its apply: the
// AmatrixClassLMCuda accessors such an apply uses (getDataPointer_d,
// getRowSize, getColSize, getDimension; AmatrixClassLMCuda.h:130-349) over a
// column-major m x n device block with leading dimension getLeadingDimension()
// (= m), and the HIP stream (as void*) its work must go to.  A view of a
// library buffer: valid during the apply only.
template <typename T = double>
class DeviceMatrix {
 public:
  DeviceMatrix(void* data, long m, long n, long ld, void* stream)
      : d_(static_cast<T*>(data)), m_(m), n_(n), ld_(ld), stream_(stream) {}
  T* getDataPointer_d() const { return d_; }
  long getRowSize() const { return m_; }
  long getColSize() const { return n_; }
  long getLeadingDimension() const { return ld_; }
  size_t getDimension() const { return (size_t)m_ * (size_t)n_; }
  void* stream() const { return stream_; }

 private:
  T* d_;
  long m_, n_, ld_;
  void* stream_;
};

// A caller's operator as code, the reference's contract
// Otype::apply(Atype& in, Atype& out, alpha, beta): out = alpha A in + beta out
// (LanczosCpolyOperatorLMCuda.h:236,245, RayleighChebyshevLMCuda.h:2135,2235).
// UserOp provides apply(DeviceMatrix<T>& in, DeviceMatrix<T>& out, double
// alpha, double beta) that enqueues its work on in.stream() (or finishes it
// before returning); T is double or std::complex<double> (Hermitian).  An
// exception from it fails the solve with std::runtime_error.  The user
// operator must outlive this adapter.
template <class UserOp, typename T = double>
class CallbackOperator {
 public:
  CallbackOperator(Context& ctx, long n, UserOp& op) : ctx_(&ctx), op_(&op), n_(n) {
    static_assert(std::is_same<T, double>::value || std::is_same<T, std::complex<double>>::value,
                  "a caller's operator is real or complex double");
    check(rcg_operator_callback_create(ctx.handle(), n,
                                       std::is_same<T, double>::value ? RCG_F64 : RCG_C128,
                                       &CallbackOperator::trampoline, this, &h_));
  }
  ~CallbackOperator() { rcg_operator_destroy(h_); }
  CallbackOperator(const CallbackOperator&) = delete;
  CallbackOperator& operator=(const CallbackOperator&) = delete;
  template <class Atype>
  void apply(Atype& in, Atype& out, double alpha, double beta) {
    check(rcg_operator_apply(h_, in.handle(), out.handle(), alpha, beta, in.getColSize()));
  }
  rcg_operator handle() const { return h_; }
  long getDimension() const { return n_; }
  Context& context() const { return *ctx_; }

 private:
  static int trampoline(void* user, const void* in, void* out, int64_t n, int64_t cols, int64_t ld,
                        double alpha, double beta, void* stream) {
    try {
      DeviceMatrix<T> a(const_cast<void*>(in), (long)n, (long)cols, (long)ld, stream);
      DeviceMatrix<T> b(out, (long)n, (long)cols, (long)ld, stream);
      static_cast<CallbackOperator*>(user)->op_->apply(a, b, alpha, beta);
      return 0;
    } catch (...) {
      return 1;
    }
  }
  Context* ctx_;
  UserOp* op_;
  rcg_operator h_ = nullptr;
  long n_ = 0;
};

// A dense real symmetric operator (column-major n x n, leading dimension lda).
class DenseOperator {
 public:
  DenseOperator(Context& ctx, long n, const double* A, long lda) : ctx_(&ctx), n_(n) {
    check(rcg_operator_dense_create(ctx.handle(), n, A, lda, RCG_MEM_HOST, &h_));
  }
  ~DenseOperator() { rcg_operator_destroy(h_); }
  DenseOperator(const DenseOperator&) = delete;
  DenseOperator& operator=(const DenseOperator&) = delete;
  template <class Atype>
  void apply(Atype& in, Atype& out, double alpha, double beta) {
    check(rcg_operator_apply(h_, in.handle(), out.handle(), alpha, beta, in.getColSize()));
  }
  rcg_operator handle() const { return h_; }
  long getDimension() const { return n_; }
  Context& context() const { return *ctx_; }

 private:
  Context* ctx_;
  rcg_operator h_ = nullptr;
  long n_ = 0;
};

// ArandomizeOpClassLMCuda role (ArandomizeOpClassLMCuda.h:43-91), seeded.
class ArandomizeOpClass {
 public:
  explicit ArandomizeOpClass(uint64_t seed = 0) : seed(seed) {}
  template <class Atype>
  void randomize(Atype& M) {
    check(rcg_block_randomize(M.handle(), 0, M.getColSize(), seed, &counter));
  }
  uint64_t seed;
  uint64_t counter = 0;
};

// RC_Types.h:37
struct RC_Types {
  enum StopCondition { DEFAULT, COMBINATION, EIGENVALUE_ONLY, RESIDUAL_ONLY };
};

template <class Atype, class Vtype, class Otype, class VRandomizeOpType, typename Dtype = double>
class RayleighChebyshevLM {
 public:
  RayleighChebyshevLM() { p_ = rcg_solve_params{}; p_.stop = RCG_STOP_COMBINATION; }

  // setters (RayleighChebyshevLM.h:281-440)
  void setStopCondition(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::toupper(c); });
    if (s == "COMBINATION" || s == "DEFAULT") p_.stop = RCG_STOP_COMBINATION;
    else if (s == "EIGENVALUE_ONLY") p_.stop = RCG_STOP_EIGENVALUE_ONLY;
    else if (s == "RESIDUAL_ONLY") p_.stop = RCG_STOP_RESIDUAL_ONLY;
    else
      throw std::runtime_error(
          "\nRayleighChebyshevLM Error : Stopping condition type specified not"
          "\none of DEFAULT, COMBINATION, EIGENVALUE_ONLY,or RESIDUAL_ONLY."
          "\nOffending specification : " + s + "\n");
  }
  void setStopCondition(RC_Types::StopCondition c) {
    p_.stop = (c == RC_Types::DEFAULT) ? RCG_STOP_COMBINATION : (int)c;
  }
  void setMaxMinTolerance(double v = 1.0e-6) { p_.maxmin_tol = v; }
  void setMaxInnerLoopCount(long v) { p_.max_inner = v; }
  void resetMaxInnerLoopCount() { p_.max_inner = 0; }
  void setFixedIterationCount(bool v) { p_.fixed_iteration = v; }
  void clearFixedIteratonCount() { p_.fixed_iteration = 0; }
  void setNonRandomStartFlag(bool v = true) { nonRandom_ = v; }
  void clearNonRandomStartFlag() { nonRandom_ = false; }
  void setVerboseFlag(bool v = true) { p_.verbose = v; }
  void clearVerboseFlag() { p_.verbose = 0; }
  void setMinIntervalPolyDegreeMax(long v) { p_.deg_max = v; }
  void setFilterRepetitionCount(long v) { p_.rep = v; }
  void setMaxSpectralBoundsIter(long v) { p_.lanczos_itmax = v; }
  // Beyond the reference: FP32 filter (mixed precision) and a final FP64
  // Rayleigh-Ritz over all found vectors.
  void setFilterPrecision(int dtype) { p_.filter_dtype = dtype; }
  void setRefinement(bool v) { p_.refine = v; }

  // getters
  double getMinEigValueEst() const { return r_.min_est; }
  double getMaxEigValueEst() const { return r_.max_est; }
  double getGuardEigenvalue() const { return r_.guard; }
  long getSpectralBoundsIterCount() const { return (long)r_.lanczos_iters; }
  std::vector<double> getEigVectorResiduals() const { return res_; }
  std::map<std::string, long> getCountData() const {
    return {{"ortho", (long)r_.count_ortho}, {"OpApply", (long)r_.count_apply},
            {"eigenvalue", (long)r_.count_eig}};
  }
  std::map<std::string, double> getFinalData() const {
    return {{"maxResidual", r_.final_maxres}, {"maxEigValueDifference", r_.final_maxdiff},
            {"maxRelEigValueGap", r_.final_maxgap}};
  }
  const rcg_solve_result& getResult() const { return r_; }

  // getMinEigenSystem (RayleighChebyshevLM.h:712-738)
  long getMinEigenSystem(long eigCount, double subspaceTol, long inc, long buf, Vtype& vStart,
                         Otype& oP, VRandomizeOpType& randOp, std::vector<double>& eigValues,
                         Atype& eigVectors) {
    (void)vStart;
    return solve(eigCount, subspaceTol, inc, buf, oP, randOp, eigValues, eigVectors, false, 0.0,
                 false, 0.0, 0.0);
  }
  // getMinEigenSystem with given bounds (RayleighChebyshevLM.h:693-708)
  long getMinEigenSystem(long eigCount, double minEigValueEst, double maxEigValueBound,
                         double subspaceTol, long inc, long buf, Vtype& vStart, Otype& oP,
                         VRandomizeOpType& randOp, std::vector<double>& eigValues,
                         Atype& eigVectors) {
    (void)vStart;
    return solve(eigCount, subspaceTol, inc, buf, oP, randOp, eigValues, eigVectors, false, 0.0,
                 true, minEigValueEst, maxEigValueBound);
  }
  // The CPU template's form (RayleighChebyshevLM.h:712-738): eigenvectors
  // returned as host vectors (Vtype with a vData member, e.g. AvectorClass).
  long getMinEigenSystem(long eigCount, double subspaceTol, long inc, long buf, Vtype& vStart,
                         Otype& oP, VRandomizeOpType& randOp, std::vector<double>& eigValues,
                         std::vector<Vtype>& eigVectors) {
    const long cap = std::max(eigCount, inc + buf);
    Atype V(oP.context(), oP.getDimension(), cap);
    if (nonRandom_) {
      const long c = std::min<long>((long)eigVectors.size(), cap);
      for (long j = 0; j < c; ++j)
        for (long i = 0; i < oP.getDimension(); ++i) V(i, j) = eigVectors[j].vData[i];
      V.host_to_device();
      V.setValidColumns(c);
    }
    const long f = getMinEigenSystem(eigCount, subspaceTol, inc, buf, vStart, oP, randOp,
                                     eigValues, V);
    V.device_to_host();
    eigVectors.assign(f, vStart);
    for (long j = 0; j < f; ++j)
      for (long i = 0; i < oP.getDimension(); ++i) eigVectors[j].vData[i] = V(i, j);
    return f;
  }
  // getMinIntervalEigenSystem (RayleighChebyshevLM.h:745-782)
  long getMinIntervalEigenSystem(double lambdaMax, double subspaceTol, long inc, long buf,
                                 long maxEigensystemDim, Vtype& vStart, Otype& oP,
                                 VRandomizeOpType& randOp, std::vector<double>& eigValues,
                                 Atype& eigVectors) {
    (void)vStart;
    return solve(maxEigensystemDim, subspaceTol, inc, buf, oP, randOp, eigValues, eigVectors,
                 true, lambdaMax, false, 0.0, 0.0);
  }
  // getMinIntervalEigenSystem with the caller's spectral bounds, no Lanczos
  // (RayleighChebyshev.h:812-821).
  long getMinIntervalEigenSystem(double minEigValueEst, double lambdaMax, double maxEigValueBound,
                                 double subspaceTol, long inc, long buf, long maxEigensystemDim,
                                 Vtype& vStart, Otype& oP, VRandomizeOpType& randOp,
                                 std::vector<double>& eigValues, Atype& eigVectors) {
    (void)vStart;
    return solve(maxEigensystemDim, subspaceTol, inc, buf, oP, randOp, eigValues, eigVectors,
                 true, lambdaMax, true, minEigValueEst, maxEigValueBound);
  }

 protected:
  bool nonRandomStart() const { return nonRandom_; }

 private:
  long solve(long k, double tol, long inc, long buf, Otype& oP, VRandomizeOpType& randOp,
             std::vector<double>& eigValues, Atype& eigVectors, bool interval, double lambdaMax,
             bool bounds, double mn, double mx) {
    rcg_solve_params p = p_;
    p.k = k;
    p.inc = inc;
    p.buf = buf;
    p.tol = tol;
    p.seed = randOp.seed;
    p.interval = interval;
    p.lambda_max = lambdaMax;
    p.use_bounds = bounds;
    p.min_eig = mn;
    p.max_eig = mx;
    const long cap = std::max(k, inc + buf);
    if (eigVectors.getColSize() < cap) eigVectors.resize(cap);
    p.nonrandom_start = nonRandom_ ? (int)std::min(eigVectors.validColumns(), eigVectors.getColSize()) : 0;
    std::vector<double> ev(cap), rr(cap);
    check(rcg_solve_lowest(oP.handle(), &p, &r_, ev.data(), rr.data(), cap, eigVectors.handle(),
                           nullptr, 0));
    const long f = (long)r_.found;
    randOp.counter = r_.rng_counter;
    eigValues.assign(ev.begin(), ev.begin() + f);
    res_.assign(rr.begin(), rr.begin() + f);
    eigVectors.setValidColumns(f);
    return f;
  }

  rcg_solve_params p_;
  rcg_solve_result r_{};
  std::vector<double> res_;
  bool nonRandom_ = false;
};

// The older RayleighChebyshev template (RayleighChebyshev.h:210-821): the
// same solves, with eigenvectors exchanged as host vectors (Vtype with a vData
// member, e.g. AvectorClass<double>) in a std::vector.  With the non-random
// start flag the vectors passed in are the start block.  Each call moves the
// block between host and device once.
template <class Vtype, class Otype, class VRandomizeOpType, typename Dtype = double>
class RayleighChebyshev
    : public RayleighChebyshevLM<AmatrixClass<Dtype>, Vtype, Otype, VRandomizeOpType, Dtype> {
  using Base = RayleighChebyshevLM<AmatrixClass<Dtype>, Vtype, Otype, VRandomizeOpType, Dtype>;

 public:
  // RayleighChebyshev.h:728-755
  long getMinEigenSystem(long eigCount, double subspaceTol, long inc, long buf, Vtype& vStart,
                         Otype& oP, VRandomizeOpType& randOp, std::vector<double>& eigValues,
                         std::vector<Vtype>& eigVectors) {
    return host_solve(std::max(eigCount, inc + buf), vStart, oP, eigVectors,
                      [&](AmatrixClass<Dtype>& V) {
                        return Base::getMinEigenSystem(eigCount, subspaceTol, inc, buf, vStart, oP,
                                                       randOp, eigValues, V);
                      });
  }
  // RayleighChebyshev.h:709-724
  long getMinEigenSystem(long eigCount, double minEigValueEst, double maxEigValueBound,
                         double subspaceTol, long inc, long buf, Vtype& vStart, Otype& oP,
                         VRandomizeOpType& randOp, std::vector<double>& eigValues,
                         std::vector<Vtype>& eigVectors) {
    return host_solve(std::max(eigCount, inc + buf), vStart, oP, eigVectors,
                      [&](AmatrixClass<Dtype>& V) {
                        return Base::getMinEigenSystem(eigCount, minEigValueEst, maxEigValueBound,
                                                       subspaceTol, inc, buf, vStart, oP, randOp,
                                                       eigValues, V);
                      });
  }
  // RayleighChebyshev.h:761-805
  long getMinIntervalEigenSystem(double lambdaMax, double subspaceTol, long inc, long buf,
                                 long maxEigensystemDim, Vtype& vStart, Otype& oP,
                                 VRandomizeOpType& randOp, std::vector<double>& eigValues,
                                 std::vector<Vtype>& eigVectors) {
    return host_solve(std::max(maxEigensystemDim, inc + buf), vStart, oP, eigVectors,
                      [&](AmatrixClass<Dtype>& V) {
                        return Base::getMinIntervalEigenSystem(lambdaMax, subspaceTol, inc, buf,
                                                               maxEigensystemDim, vStart, oP,
                                                               randOp, eigValues, V);
                      });
  }
  // RayleighChebyshev.h:812-821
  long getMinIntervalEigenSystem(double minEigValueEst, double lambdaMax, double maxEigValueBound,
                                 double subspaceTol, long inc, long buf, long maxEigensystemDim,
                                 Vtype& vStart, Otype& oP, VRandomizeOpType& randOp,
                                 std::vector<double>& eigValues, std::vector<Vtype>& eigVectors) {
    return host_solve(std::max(maxEigensystemDim, inc + buf), vStart, oP, eigVectors,
                      [&](AmatrixClass<Dtype>& V) {
                        return Base::getMinIntervalEigenSystem(
                            minEigValueEst, lambdaMax, maxEigValueBound, subspaceTol, inc, buf,
                            maxEigensystemDim, vStart, oP, randOp, eigValues, V);
                      });
  }

 private:
  template <class F>
  long host_solve(long cap, Vtype& vStart, Otype& oP, std::vector<Vtype>& eigVectors, F&& call) {
    const long n = oP.getDimension();
    AmatrixClass<Dtype> V(oP.context(), n, cap);
    if (this->nonRandomStart()) {
      const long c = std::min<long>((long)eigVectors.size(), cap);
      for (long j = 0; j < c; ++j)
        for (long i = 0; i < n; ++i) V(i, j) = eigVectors[j].vData[i];
      V.host_to_device();
      V.setValidColumns(c);
    }
    const long f = call(V);
    V.device_to_host();
    eigVectors.assign(f, vStart);
    for (long j = 0; j < f; ++j) {
      eigVectors[j].vData.resize(n);
      for (long i = 0; i < n; ++i) eigVectors[j].vData[i] = V(i, j);
    }
    return f;
  }
};

}  // namespace rcg
