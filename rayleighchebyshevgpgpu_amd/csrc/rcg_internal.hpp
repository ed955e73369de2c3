// rcg_internal.hpp -- shared declarations of the HIP kernels' host launchers
// and the library's internal handle types.
//
// Device layout of a block (the reference's AmatrixClassLMCuda, AM:12-359, is
// column-major with ld = N): here an N x cols block is stored row-interleaved,
// element (i, c) at data[i*ld + c], ld a multiple of 4 elements.  A CSR row of
// the operator then gathers whole, contiguous rows of the block.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace rcg {

enum DType : int { F64 = 0, F32 = 1, C128 = 2, C64 = 3 };

inline size_t dtype_size(int dt) {
  switch (dt) {
    case F64: return 8;
    case F32: return 4;
    case C128: return 16;
    case C64: return 8;
  }
  return 0;
}

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define RCG_HIP(call)                                                                 \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw ::rcg::Error(2, std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                                __FILE__ + ":" + std::to_string(__LINE__) + " (" #call ")"); \
  } while (0)

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// The calling thread's message for rcg_get_last_error.
void set_last_error(const std::string& msg);

// -------------------------------------------------------------- launchers
// All pointers are device pointers; every launch goes on `st`.

// Order in which the CSR kernels visit row blocks.  The rows computed and
// their arithmetic do not depend on it; it only changes which rows are in
// flight together (cache locality).  order == nullptr: contiguous blocks.
struct CsrSched {
  const int* order = nullptr;  // device: logical block -> row block
  int64_t rows_per_blk = 0;    // rows per block when `order` is set
  int64_t nblk = 0;            // number of row blocks when `order` is set
  int remap = 1;               // 1: each XCD takes a contiguous range of logical blocks
};

// ---- batched filter steps over many CSR operators (the serving engine)
// One request: an operator and its ping-pong blocks.  dtype of val/buf/R/kap
// is double (plain FP64 filter) or float (FP32 correction steps).
struct BatchDesc {
  const int* rowptr;
  const int* col;
  const void* val;
  void* buf[2];
  const void* R;    // correction source block (FP32 steps only)
  const void* kap;  // correction coefficient rows, row j at kap + j*ld
  int64_t n;
  int ld;           // real elements per row
  int ncv;          // 16-byte column vectors per row
  int tail;         // columns of the call in the last vector (1..elements per vector)
  int m;            // filter degree
  int rows_per_blk;
  double g1, g2;
};
// Largest row block of a batched task (the kernel stages its row pointers
// in LDS).
constexpr int kBatchMaxRows = 512;
struct BatchTask {
  int req;
  int blk;
};
// Launch class (lanes per row, one vector per lane) of a request with ncv
// 16-byte vectors per row; requests of one class share a launch.  Tasks of
// class cls span batch_rows(cls) rows.
int batch_class(int ncv);
int batch_rows(int cls);
// Step j of every task's request (all tasks of one batch_class); src: FP32
// correction steps (j >= 2).  Descriptors and tasks live in device memory.
void launch_cheb_batch(const BatchDesc* descs, const BatchTask* tasks, int ntasks, int j,
                       bool f32, bool src, bool cx, int cls, hipStream_t st);

// out = alpha*A*X + beta*out over `ncols` columns (ld shared by X and out).
// Complex operators (cx): val holds interleaved (re, im) pairs, the blocks
// interleave each complex column as two real columns, and ncols counts real
// columns.
template <typename T>
void launch_csr_apply(int64_t n, const int* rowptr, const int* col, const T* val, const T* X,
                      T* out, int ld, int ncols, T alpha, T beta, hipStream_t st,
                      const CsrSched* sched = nullptr, bool cx = false);

// One three-term Chebyshev step of the scaled C polynomial:
//   first: out = s * g2*(g1*A*X1 + X1)
//   else : out = s * (g2*(g1*A*X1 + X1) - X0)
// `out` may alias X0 (each row reads its own X0 entries before writing them).
// `panel` > 0 splits the columns into launches of that width (a multiple of 4).
template <typename T>
void launch_cheb_step(int64_t n, const int* rowptr, const int* col, const T* val, const T* X1,
                      const T* X0, T* out, int ld, int ncols, T g1, T g2, T s, bool first,
                      hipStream_t st, int panel = 0, const CsrSched* sched = nullptr,
                      bool cx = false);

// The mixed filter's FP64 ends fused into its kernels: the start (below) and
// the recombination V <- s_end (um[c] V + E_m) computed in the last FP32
// step's epilogue (launch_cheb_step_src_end; single-operator filters only:
// in the batched steps a kernel of its own for the ending requests served
// 1588 -> 1580 eigenpairs/s).
struct MixedArgs {
  const double* cp = nullptr;  // theta_c at cp[c], the Ritz flag at cp[nc + c]
  float* e0 = nullptr;
  float* e1 = nullptr;
  float* r = nullptr;
  double e1c = 0.0, g1 = 0.0, g2 = 0.0;
  int nc = 0;  // real columns of the call
  double* v_end = nullptr;       // the recombination's V (same element layout as E)
  const double* um = nullptr;    // u_m(theta_c) per column
  double s_end = 0.0;            // 1 / (m + 1)
};

// Mixed-precision filter.  The FP64 filter output s*X_m of X_0 = v is split
// per column as X_j = u_j(theta_c) v + E_j, u_j the scalar recurrence of the
// step at the Ritz value theta_c; E_j (FP32) then obeys
//   E_j = g2*(g1*A*E_{j-1} + E_{j-1}) - E_{j-2} + kap_j[c] * R,
//   kap_j[c] = g1*g2*u_{j-1}(theta_c), R = A v - theta_c v (FP64 -> FP32),
// so FP32 rounding is relative to the residual, not to v.
void launch_cheb_step_src(int64_t n, const int* rowptr, const int* col, const float* val,
                          const float* X1, const float* X0, float* out, const float* R,
                          const float* kap, int ld, int ncols, float g1, float g2, float s,
                          hipStream_t st, int panel = 0, const CsrSched* sched = nullptr,
                          bool cx = false);
// The last correction step with the recombination in its epilogue
// (end.v_end, end.um, end.s_end over the same ncols columns): V is written,
// X_m is not.  Bit-identical to launch_cheb_step_src + launch_mixed_combine.
void launch_cheb_step_src_end(int64_t n, const int* rowptr, const int* col, const float* val,
                              const float* X1, const float* X0, const float* R, const float* kap,
                              int ld, int ncols, float g1, float g2, const MixedArgs& end,
                              hipStream_t st, int panel = 0, const CsrSched* sched = nullptr,
                              bool cx = false);
// E0, E1, R from V and AV (see the kernel); cp[c] = theta_c, cp[nc+c] = 1 for
// Ritz columns, 0 for plain ones (E_j = X_j, R = 0).
// W = A V and the start above in one pass: the apply's epilogue turns each
// (v, w) pair into the E0, E1, R entries, so W is never stored.  The results
// are bit-identical to launch_csr_apply followed by launch_mixed_start.
void launch_csr_apply_start(int64_t n, const int* rowptr, const int* col, const double* val,
                            const double* V, int ld, int ncols, const MixedArgs& ms,
                            hipStream_t st, const CsrSched* sched = nullptr, bool cx = false);
void launch_mixed_start(int64_t n, const double* V, const double* AV, int ld, int nc,
                        const double* cp, double e1, double g1, double g2, float* E0, float* E1,
                        float* R, hipStream_t st);
// V <- s * (um[c] V + E).
void launch_mixed_combine(int64_t n, double* V, int ld, int nc, const double* um, double s,
                          const float* E, hipStream_t st);
void launch_to_f32(int64_t n, const double* x, float* y, hipStream_t st);

// Dense symmetric operator A (n x n, lda; row-major = column-major): one
// step on a row-interleaved block, mode as the CSR kernels:
//   0 apply: out = p0 A X; 1: out = p0 A X + p1 out;
//   2 first: out = p2 p1 (p0 A X + X); 3: out = p2 (p1 (p0 A X + X) - Z);
//   4 source: mode 3 + kap[c] R.  MFMA FP64 (T = double) or FP32.
enum StepMode : int { kStepApply = 0, kStepApplyBeta = 1, kStepFirst = 2, kStepCheb = 3, kStepSrc = 4 };
template <typename T>
void launch_dense_step(int64_t n, const T* A, int64_t lda, const T* X, const T* Z, T* out, int ld,
                       int ncols, T p0, T p1, T p2, int mode, const T* R, const T* kap,
                       hipStream_t st);

// G (bx x by, column-major, ldg = bx) = X^H Y over n rows; ws holds partials.
// X == Y (same pointer and ld) computes the upper 64x64 tiles and mirrors
// them.  lower_only (bx == by): the caller reads only G's lower triangle,
// whose tiles are the only ones multiplied (FP64).
template <typename T>
void launch_gram(int64_t n, const T* X, int ldx, int bx, const T* Y, int ldy, int by, T* G,
                 T* ws, size_t ws_elems, hipStream_t st, bool lower_only = false);
// G = X^T X and, right after it (G + b*b), X^T Y, b <= 64 columns, in one
// launch and one reduction: the same values as two launch_gram calls.
// Needs 2 gram_ws_elems(n, b, b) of workspace.
void launch_gram_pair(int64_t n, const double* X, int ldx, const double* Y, int ldy, int b,
                      double* G, double* ws, size_t ws_elems, hipStream_t st);
template <typename T>
size_t gram_ws_elems(int64_t n, int bx, int by);
// Bound of gram_ws_elems over all widths <= bmax.
size_t gram_ws_bound(int64_t n, int bmax);

// out = alpha * X*C + beta * Z, X: n x bx (ldx), C: bx x by column-major (ldc),
// Z/out: n x by (ldo).  out may alias Z, not X.
template <typename T>
void launch_tall_gemm(int64_t n, const T* X, int ldx, int bx, const T* C, int ldc, int by,
                      T alpha, T beta, const T* Z, T* out, int ldo, hipStream_t st);

// sums[c] = sum_i |X[i,c]|^2, c < nc.
template <typename T>
void launch_col_sumsq(int64_t n, const T* X, int ld, int nc, double* sums, double* ws,
                      size_t ws_elems, hipStream_t st);
// sums[c] = sum_i |theta[c] V[i,c] - W[i,c]|^2, c < nc.
template <typename T>
void launch_residual_sumsq(int64_t n, const T* V, const T* W, int ld, int nc, const double* theta,
                           double* sums, double* ws, size_t ws_elems, hipStream_t st);
size_t colred_ws_elems(int64_t n, int nc);
// sums[c] = sum_i |theta[c] V[i,c] - W[i,c]|^2 and sums[nc + c] = sum_i |V[i,c]|^2,
// c < nc (workspace for 2 nc columns).
template <typename T>
void launch_resid_norm_sumsq(int64_t n, const T* V, const T* W, int ld, int nc,
                             const double* theta, double* sums, double* ws, size_t ws_elems,
                             hipStream_t st);
// The same with theta (nc <= kThetaArgs values) passed in the kernel's
// argument block instead of device memory (no host-to-device copy).
constexpr int kThetaArgs = 128;
struct ThetaArgs {
  double v[kThetaArgs];
#ifdef __HIPCC__
  __host__ __device__
#endif
  double operator[](int c) const { return v[c]; }
};
void launch_resid_norm_sumsq_args(int64_t n, const double* V, const double* W, int ld, int nc,
                                  const ThetaArgs& theta, double* sums, double* ws,
                                  size_t ws_elems, hipStream_t st);

// Caller operators (column-major blocks) around the row-interleaved solver
// blocks: `cols` entries per row (cw = 2: complex entries of two doubles),
// the rows layout with ldr_real doubles per row, the column-major one with
// ldc entries per column.  to_cols: rows -> column-major, else the reverse.
void launch_layout(int64_t n, int cols, int cw, double* rows, int64_t ldr_real, double* colmaj,
                   int64_t ldc, bool to_cols, hipStream_t st);
// out = p2 (p1 (p0 AX + X) - Z) (rows layout; AX column-major, Z optional
// and may alias out): a Chebyshev step around a caller operator's product.
void launch_cheb_combine(int64_t n, int cols, int cw, const double* AX, int64_t ldc,
                         const double* X, const double* Z, double* out, int64_t ldr_real,
                         double p0, double p1, double p2, hipStream_t st);
// alpha_k = Re sum conj(w) v, alpha_im_k = Im of it (interleaved complex n-vectors)
void launch_lz_dot_c(int64_t n, const double* w, const double* v, double* alpha_k,
                     double* alpha_im_k, double* ws, size_t ws_elems, hipStream_t st);

// out[c] = sqrt(|sums[cw c] + ... + sums[cw c + cw - 1]|), c < count: the
// residual norm of each (complex, cw = 2) column from its real columns' sums.
void launch_group_sqrt(const double* sums, int cw, int count, double* out, hipStream_t st);

// X[i,c] *= 1/sqrt(sums[c]), c < nc.
template <typename T>
void launch_normalize_cols(int64_t n, T* X, int ld, int nc, const double* sums, hipStream_t st);
// dst[i, d0+j] = src[i, s0+j], j < cnt.
template <typename T>
void launch_copy_cols(int64_t n, const T* src, int lds, int s0, T* dst, int ldd, int d0, int cnt,
                      hipStream_t st);
// X[i, c0+j] = U[-1,1) draw (counter + j*n + i) of stream `seed` (complex: two
// draws per element, counter + 2(j*n + i) + {0,1}).
template <typename T>
void launch_randomize_cols(int64_t n, T* X, int ld, int c0, int cnt, uint64_t seed,
                           uint64_t counter, hipStream_t st);
template <typename T>
void launch_fill_cols(int64_t n, T* X, int ld, int c0, int cnt, T value, hipStream_t st);

// ---- single-vector Lanczos (LanczosMaxMinFinder.h:248-544), FP64 only.
// Scalars alpha[k], beta[k] live in device memory.
void launch_lz_start(int64_t n, double* v, double* w, double* ws, size_t ws_elems,
                     double* scratch, hipStream_t st);
void launch_lz_swap(int64_t n, double* v, double* w, const double* beta_prev, hipStream_t st);
void launch_lz_apply_dot(int64_t n, const int* rowptr, const int* col, const double* val,
                         const double* w, double* v, double* alpha_k, double* ws,
                         size_t ws_elems, hipStream_t st);
void launch_lz_axpy_norm(int64_t n, const double* w, double* v, const double* alpha_k,
                         double* beta_k, double* ws, size_t ws_elems, hipStream_t st);
size_t lz_ws_elems(int64_t n);
// One fused Lanczos step for real CSR operators: reads (v, w), writes (vn, wn),
// alpha_k and beta_k; *cnt must be zero before the first call (each kernel
// leaves it zero).  beta_prev == nullptr on the first step.
void launch_lz_step(int64_t n, const int* rowptr, const int* col, const double* val,
                    const double* v, const double* w, double* vn, double* wn,
                    const double* beta_prev, double* alpha_k, double* beta_k, unsigned* cnt,
                    double* ws, size_t ws_elems, hipStream_t st);
// *out = w . v (two-phase, fixed order)
void launch_lz_dot(int64_t n, const double* w, const double* v, double* out, double* ws,
                   size_t ws_elems, hipStream_t st);
// complex (interleaved) v += A w; alpha_k = Re sum conj(w) v, alpha_im_k = Im.
void launch_lz_apply_dot_c(int64_t n, const int* rowptr, const int* col, const double* val,
                           const double* w, double* v, double* alpha_k, double* alpha_im_k,
                           double* ws, size_t ws_elems, hipStream_t st);

// ---- complex helpers
// complex columns [c0, c0+cnt) of an interleaved block: (re, im) of row i,
// column c0+j from draws counter + 2(j*n + i) + {0, 1}.
void launch_randomize_cplx(int64_t n, double* X, int ld, int c0, int cnt, uint64_t seed,
                           uint64_t counter, hipStream_t st);
// s[2c] = s[2c+1] = s[2c] + s[2c+1], c < nc.
void launch_pair_sums(double* s, int nc, hipStream_t st);
// real Gram of interleaved blocks (2f x 2b) -> real embedding of Y^H V, in place.
void launch_herm_embed(double* G, int f, int b, hipStream_t st);


}  // namespace rcg
