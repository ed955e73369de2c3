/*
 * rcg_c_api.h -- C ABI of librcg.so, the MI355X (gfx950) Rayleigh-Chebyshev
 * lowest-eigenpair solver.
 *
 * Plain pointers, sizes and opaque handles; no HIP or C++ types cross the
 * boundary (a stream is passed as void*).  Every function returns an int
 * status (0 = RCG_OK); rcg_get_last_error() returns the message of the last
 * failure on the calling thread.  No function exits the process (the
 * reference aborts through cudaErrchk.h:8-50).
 *
 * Each entry point names the reference interface it replaces
 * (file:line in almaeder/RayleighChebyshevGPGPU).
 *
 * Device layout: an N x cols block is row-interleaved, element (i, c) at
 * ((T*)data)[i*ld + c] with ld a multiple of 4; host arrays crossing the
 * boundary are column-major like the reference's AmatrixClassLMCuda
 * (AmatrixClassLMCuda.h:333-341).
 */
#ifndef RCG_C_API_H_
#define RCG_C_API_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rcg_context_s* rcg_context;
typedef struct rcg_operator_s* rcg_operator;
typedef struct rcg_block_s* rcg_block;
typedef struct rcg_server_s* rcg_server;

enum rcg_status {
  RCG_OK = 0,
  RCG_ERR_INVALID = 1,       /* bad argument (std::runtime_error in the reference) */
  RCG_ERR_HIP = 2,           /* HIP runtime failure */
  RCG_ERR_ALLOC = 3,         /* device or workspace allocation failed */
  RCG_ERR_NOT_HERMITIAN = 4, /* std::invalid_argument, LanczosMaxMinFinder.h:350-356 */
  RCG_ERR_NUMERIC = 5,       /* breakdown the solve could not recover from */
  RCG_ERR_UNSUPPORTED = 6
};

enum rcg_dtype { RCG_F64 = 0, RCG_F32 = 1, RCG_C128 = 2, RCG_C64 = 3 };

/* RC_Types.h:37 StopCondition */
enum rcg_stop {
  RCG_STOP_DEFAULT = 0,
  RCG_STOP_COMBINATION = 1,
  RCG_STOP_EIGENVALUE_ONLY = 2,
  RCG_STOP_RESIDUAL_ONLY = 3
};

/* Operator memory flags for rcg_operator_csr_create. */
enum rcg_mem { RCG_MEM_HOST = 0, RCG_MEM_DEVICE = 1 };

/* ------------------------------------------------------------ errors */
/* Replaces cudaErrchk/cublasErrchk/cusolverErrchk/cusparseErrchk
 * (cudaErrchk.h:8-50): copy the last error message (NUL-terminated). */
int rcg_get_last_error(char* buf, size_t len);
/* Library version string, e.g. "rcg 0.1 gfx950". */
const char* rcg_version(void);
/* Number of HIP devices visible to the process. */
int rcg_device_count(int* count);

/* ----------------------------------------------------------- context */
/* Device + stream binding for every handle made from it.  hip_stream may be
 * NULL (the library creates a default-priority stream).  The reference runs
 * every launch on the legacy default stream (RayleighChebyshevLMCuda.h). */
int rcg_context_create(int device, void* hip_stream, rcg_context* out);
int rcg_context_destroy(rcg_context ctx);
int rcg_context_synchronize(rcg_context ctx);
/* No counterpart in the reference.  Solve buffers, workspaces and pinned
 * staging are cached by the library after use (hipFree would synchronise
 * the device and stall concurrent solves); at most 24 GiB per device are
 * kept.  This frees every cached buffer now. */
int rcg_memory_trim(void);

/* ------------------------------------------------------------ blocks */
/* AmatrixClassLMCuda(m, n) / resize / create_cuda_memory
 * (AmatrixClassLMCuda.h:21-80, 251-266): an N x cols device block,
 * zero-filled; dtype RCG_F64, RCG_F32 or RCG_C128 (ld counts elements). */
int rcg_block_create(rcg_context ctx, int64_t n, int64_t cols, int dtype, rcg_block* out);
int rcg_block_destroy(rcg_block b);
/* host_to_device (AmatrixClassLMCuda.h:109-117): host column-major, leading
 * dimension ld_host >= n, first `cols` columns of the block. */
int rcg_block_upload(rcg_block b, const void* host, int64_t ld_host, int64_t cols);
/* device_to_host (AmatrixClassLMCuda.h:119-127). */
int rcg_block_download(rcg_block b, void* host, int64_t ld_host, int64_t cols);
/* Device-to-device block copy (the cudaMemcpy D2D calls at
 * LanczosCpolyOperatorLMCuda.h:216-263 and RayleighChebyshevLMCuda.h:2232). */
int rcg_block_copy(rcg_block dst, rcg_block src);
/* ArandomizeOpClassLMCuda::randomize(AmatrixClassLMCuda&)
 * (ArandomizeOpClassLMCuda.h:74-91): columns [col0, col0+ncols) get U[-1,1)
 * draws counter + j*n + i of stream `seed` (two draws per complex element);
 * *counter advances by the draws consumed. */
int rcg_block_randomize(rcg_block b, int64_t col0, int64_t ncols, uint64_t seed,
                        uint64_t* counter);
/* Shape and device pointer of a block. */
int rcg_block_info(rcg_block b, int64_t* n, int64_t* cols, int64_t* ld, int* dtype,
                   void** device_ptr);

/* --------------------------------------------------------- operators */
/* The user operator of the reference (not shipped there; its contract is
 * Otype::apply(Atype& in, Atype& out, alpha, beta), LanczosCpolyOperatorLMCuda.h
 * :236,245, RayleighChebyshevLMCuda.h:2135,2235): an N x N CSR matrix with
 * int32 indices and sorted columns.  Arrays are copied to the device
 * (mem = RCG_MEM_HOST) or referenced (RCG_MEM_DEVICE: the caller keeps them
 * alive).  val_dtype: RCG_F64, or RCG_C128 for a Hermitian operator (values
 * as interleaved re, im pairs, the layout of double _Complex); the FP32 /
 * complex64 copy for mixed-precision filtering is made on demand.  A complex
 * operator takes C128 blocks in every entry point that has an operator; the
 * block-only entry points (rcg_orthonormalize, rcg_deflate, rcg_rotate) take
 * FP64 or C128 blocks, like the reference's D- and Z- library calls
 * (RayleighChebyshevLMCuda.h:2030-2047, 2139-2161). */
int rcg_operator_csr_create(rcg_context ctx, int64_t n, int64_t nnz, const int32_t* rowptr,
                            const int32_t* col, const void* val, int val_dtype, int mem,
                            rcg_operator* out);
/* No counterpart in the reference: an operator with the sparsity pattern of
 * the CSR operator `base` (its row pointers and columns, device-resident,
 * shared, not copied) and values of its own (nnz values of base's dtype,
 * copied from the host or referenced on the device), e.g. the members of a
 * parameter sweep or of a served stream over one grid.  base must outlive the
 * new operator.  The grid hint is inherited. */
int rcg_operator_csr_share(rcg_operator base, const void* val, int mem, rcg_operator* out);
/* The user operator as a dense real symmetric n x n matrix (column-major
 * with leading dimension lda; symmetric, so also row-major).  Copied to the
 * device (RCG_MEM_HOST) or referenced (RCG_MEM_DEVICE).  Its filter steps are
 * MFMA block products (FP64, or FP32 for the mixed-precision filter). */
int rcg_operator_dense_create(rcg_context ctx, int64_t n, const double* A, int64_t lda, int mem,
                              rcg_operator* out);
/* The reference's user operator as code: its contract is
 * Otype::apply(Atype& in, Atype& out, alpha, beta), out = alpha*A*in +
 * beta*out (LanczosCpolyOperatorLMCuda.h:236,245, RayleighChebyshevLMCuda.h
 * :2135,2235).  apply(user, in, out, n, cols, ld, alpha, beta, hip_stream)
 * gets n x cols column-major device blocks (the AmatrixClassLMCuda layout,
 * leading dimension ld; FP64, or interleaved (re, im) pairs for RCG_C128)
 * and must enqueue its work on hip_stream (the context's stream) or finish it
 * before returning; it returns 0, anything else fails the calling entry point
 * with RCG_ERR_INVALID.  in and out are library buffers valid for the call
 * only; apply must not keep them or call back into this library with them.
 * The operator must be Hermitian.  Its filter runs unfused: per degree one
 * copy of the block into the column-major layout, the caller's product and
 * one pass that forms the Chebyshev step from it.  Mixed-precision filtering
 * (it needs FP32 operator values), the server and rcg_residuals_noalloc take
 * CSR and dense operators only (RCG_ERR_UNSUPPORTED). */
typedef int (*rcg_apply_fn)(void* user, const void* in, void* out, int64_t n, int64_t cols,
                            int64_t ld, double alpha, double beta, void* hip_stream);
int rcg_operator_callback_create(rcg_context ctx, int64_t n, int dtype, rcg_apply_fn apply,
                                 void* user, rcg_operator* out);
/* Operator ingest.  The reference's
 * CSR index width is the includer's RAY_INT (RayleighChebyshevLMCuda.h:409ff):
 * this takes 64-bit row pointers and columns and narrows them to the int32
 * the kernels read (RCG_ERR_UNSUPPORTED beyond 2^31 - 1). */
int rcg_operator_csr64_create(rcg_context ctx, int64_t n, int64_t nnz, const int64_t* rowptr,
                              const int64_t* col, const void* val, int val_dtype,
                              rcg_operator* out);
/* Matrix Market coordinate files (real, integer, complex or pattern; general,
 * symmetric or hermitian): rows sorted, duplicates summed, the stored
 * triangle mirrored.  _info and _csr work on the host only (nnz counts the
 * expanded entries; val holds nnz doubles, or 2 nnz interleaved for C128). */
int rcg_matrix_market_info(const char* path, int64_t* n, int64_t* nnz, int* dtype);
int rcg_matrix_market_csr(const char* path, int32_t* rowptr, int32_t* col, void* val);
int rcg_operator_from_matrix_market(rcg_context ctx, const char* path, rcg_operator* out);
/* No counterpart in the reference: a locality hint for operators on a
 * structured nx x ny x nz grid whose rows are numbered x + nx*(y + ny*z).
 * The CSR kernels may then visit x-lines in (y, z) tiles; results do not
 * depend on it.  The visiting order is autotuned on first use. */
int rcg_operator_set_grid(rcg_operator op, int64_t nx, int64_t ny, int64_t nz);
int rcg_operator_destroy(rcg_operator op);
/* Otype::apply(in, out, alpha, beta): out = alpha*A*in + beta*out over the
 * first `cols` columns (blocks share n and ld; C128 blocks for a C128
 * operator). */
int rcg_operator_apply(rcg_operator op, rcg_block in, rcg_block out, double alpha, double beta,
                       int64_t cols);

/* ---------------------------------------------------- solve pieces */
/* LanczosCpolyOperatorLMCuda::apply (LanczosCpolyOperatorLMCuda.h:194-265)
 * and LanczosCpolyOperatorLM::apply (LanczosCpolyOperatorLM.h:192-304):
 * X <- P_m(A')^rep X with A' = (A + shift)/rhoB, rhoB = (lambda_max + shift)
 * / UpperX*(m), over the first `cols` columns; tmp is a block of X's shape
 * used as the second recurrence buffer.  Columns past `cols` are not
 * written.  C128 blocks with a C128 operator (complex columns). */
int rcg_cheb_filter(rcg_operator op, rcg_block X, rcg_block tmp, int64_t cols, int64_t degree,
                    int64_t repetition, double lambda_max, double shift);
/* No counterpart in the reference (it filters one operator per call): the
 * FP64 filter of rcg_cheb_filter for `count` CSR operators of one context in
 * shared launches -- one per Chebyshev step and launch class over the 64-row
 * blocks of every operator whose degree reaches that step -- as the server
 * filters its concurrent solves.  For each q: X[q] <- P_m(A_q')^1 X[q] over
 * cols[q] columns with m = degree[q], lambda_max[q], shift[q]; tmp[q] is the
 * second recurrence buffer.  Each operator's result equals rcg_cheb_filter's
 * on it, bit for bit.  Waits for the context's stream. */
int rcg_cheb_filter_batch(int64_t count, const rcg_operator* ops, const rcg_block* X,
                          const rcg_block* tmp, const int64_t* cols, const int64_t* degree,
                          const double* lambda_max, const double* shift);
/* No counterpart in the reference (its filter is FP64 only): the same filter
 * with FP32 arithmetic.  Columns with ritz[c] != 0 are split as
 * X_j = u_j(theta[c]) x + E_j (u_j the filter's scalar recurrence at the
 * Ritz value theta[c]); E_j runs in FP32 from the FP64 residual
 * A x - theta[c] x, so the FP32 rounding error scales with that residual.
 * Other columns (all when ritz is NULL) are filtered directly in FP32.  The
 * start and the final recombination are FP64; X stays FP64. */
int rcg_cheb_filter_mixed(rcg_operator op, rcg_block X, rcg_block tmp, int64_t cols,
                          int64_t degree, int64_t repetition, double lambda_max, double shift,
                          const double* theta, const int32_t* ritz);
/* RayleighChebyshevLM::orthogonalize (RayleighChebyshevLM.h:1782-1837) /
 * cuSOLVER geqrf+orgqr (RayleighChebyshevLMCuda.h:1979-2107): X <- Q with
 * X = QR, R upper triangular with positive diagonal (CholQR2, shifted when
 * the Gram matrix is numerically singular); tmp as above. */
int rcg_orthonormalize(rcg_block X, rcg_block tmp, int64_t cols);
/* RayleighChebyshevLM::OrthogonalizeAtoB (RayleighChebyshevLM.h:1844-1917):
 * X[:, :cols] <- normalize((I - Y Y^H) X) with Y = Ysrc[:, :found]
 * (classical Gram-Schmidt applied twice). */
int rcg_deflate(rcg_block X, int64_t cols, rcg_block Y, int64_t found);
/* formVtAV (RayleighChebyshevLM.h:1928-1969, RayleighChebyshevLMCuda.h
 * :2115-2163): W = A V, H = V^H W (cols x cols, column-major, host).  A C128
 * operator takes C128 blocks and H is complex: cols x cols interleaved
 * (re, im) pairs, 2 cols^2 doubles (the cublasZgemm at LMC:2139-2161). */
int rcg_project(rcg_operator op, rcg_block V, rcg_block W, int64_t cols, double* H_host);
/* V*Y rotation of createEigenVectorsAndResiduals
 * (RayleighChebyshevLMCuda.h:2200-2232): V <- normalize(V Y), Y host
 * column-major cols x cols; tmp as above.  C128 blocks take a complex Y
 * (interleaved (re, im) pairs). */
int rcg_rotate(rcg_block V, rcg_block tmp, int64_t cols, const double* Y_host);
/* computeVtVeigensystem (RayleighChebyshevLM.h:478-569,
 * RayleighChebyshevLMCuda.h:570-642): ascending eigenvalues and column
 * eigenvectors of the symmetric n x n column-major H (FP64). */
int rcg_small_eigensystem(int64_t n, const double* H, double* evals, double* evecs);
/* cuda_kernels::residuals (cuda_kernels.cu:307-382) with the apply at
 * RayleighChebyshevLMCuda.h:2235: W = A V, r_c = ||theta_c V_c - W_c|| for
 * c < count (real theta; C128 blocks for a C128 operator). */
int rcg_residuals(rcg_operator op, rcg_block V, rcg_block W, int64_t count, const double* theta,
                  double* res_host);
/* No counterpart in the reference: rcg_residuals for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results; it calls no
 * hipMalloc, hipFree or allocating routine.  Its reduction partials, Ritz
 * values and sums live in a library-owned device array of 16 MiB per device
 * (2^21 doubles), held until the process ends: 3 doubles per real column
 * (2 per complex column) plus one per real column and 512 rows must fit, else
 * RCG_ERR_INVALID.  The operator's CSR rows are visited in the default order
 * (no autotuning).  All calls on one device share that array: calls that
 * overlap on one device race.  Waits for its stream before returning. */
int rcg_residuals_noalloc(rcg_operator op, rcg_block V, rcg_block W, int64_t count,
                          const double* theta, double* res_host);
/* LanczosMaxMinFinder::getMinMaxEigenvalues (LanczosMaxMinFinder.h:248-544)
 * on the device: stop_flags bit0 = converge the maximum, bit1 = the minimum.
 * The start vector takes n draws of `seed` from *counter. */
int rcg_lanczos_bounds(rcg_operator op, double tol, int stop_flags, int64_t itmax, uint64_t seed,
                       uint64_t* counter, double* min_eig, double* max_eig, int64_t* iters);

/* ------------------------------------------------------ whole solve */
typedef struct rcg_solve_params {
  int64_t k;              /* eigCount (or maxEigensystemDim for the interval solve) */
  int64_t inc;            /* subspaceIncrementSize */
  int64_t buf;            /* bufferSize */
  double tol;             /* subspaceTol */
  uint64_t seed;          /* start-vector stream */
  int32_t stop;           /* rcg_stop */
  int32_t interval;       /* 1: getMinIntervalEigenSystem(lambda_max, ...) */
  int64_t max_inner;      /* setMaxInnerLoopCount, 0 = 10000 */
  int64_t deg_max;        /* setMinIntervalPolyDegreeMax, 0 = 200 */
  int64_t rep;            /* setFilterRepetitionCount, 0 = 1 */
  double lambda_max;      /* interval bound */
  int64_t lanczos_itmax;  /* setMaxSpectralBoundsIter, 0 = 10000 */
  double maxmin_tol;      /* setMaxMinTolerance, 0 = 1e-6 */
  int32_t fixed_iteration;/* setFixedIterationCount */
  int32_t filter_dtype;   /* RCG_F64 (plain) or RCG_F32: mixed precision -- FP32
                             Chebyshev steps on the correction X_j - u_j(theta) v
                             of each Ritz column, FP64 everywhere else, and the
                             final FP64 refinement sweep */
  int32_t refine;         /* 1: final FP64 Rayleigh-Ritz over all found vectors */
  int32_t verbose;        /* setVerboseFlag: one line per inner iteration on stdout */
  int32_t use_bounds;     /* 1: the caller's bounds min_eig, max_eig, no Lanczos:
                             getMinEigenSystem(k, min_eig, max_eig, ...) or, with
                             interval, getMinIntervalEigenSystem(min_eig, lambda_max,
                             max_eig, ...) (RayleighChebyshev.h:709-723, 812-821) */
  int32_t nonrandom_start;/* setNonRandomStartFlag: > 0 starts from the first
                             nonrandom_start columns of `evecs` (LM:957-985) */
  double min_eig;
  double max_eig;
} rcg_solve_params;

typedef struct rcg_solve_result {
  int64_t found;
  int64_t lanczos_iters;     /* getSpectralBoundsIterCount */
  double min_est, max_est;   /* getMinEigValueEst / getMaxEigValueEst */
  double guard;              /* getGuardEigenvalue */
  int64_t count_ortho, count_apply, count_eig; /* getCountData */
  double final_maxres, final_maxdiff, final_maxgap; /* getFinalData */
  int64_t trace_len;
  uint64_t rng_counter;
  int64_t filter_degree_sum; /* Chebyshev step launches (sum of starDegree * rep) */
  double t_total;            /* host wall seconds of the solve */
  double t_lanczos, t_filter, t_ortho, t_rr; /* seconds of the solve's stream per phase,
                                from hipEvents recorded where each phase begins (so a
                                phase includes the stream's idle time while the host
                                works inside it); t_ortho: start vectors, CholQR and
                                deflation; t_rr: the projection with its CholQR passes,
                                the small eigensolve, rotation, residuals and the final
                                refinement */
  double t_filter_device;    /* seconds between hipEvents around the Chebyshev step
                                launches (FP64 steps, or the FP32 correction steps) */
  double filter_bytes;       /* algorithmic bytes of those steps: CSR + 3 block passes
                                (FP64) or CSR(FP32 values) + 4 FP32 block passes */
  int64_t filter_panel;      /* autotuned column-panel width of the filter step */
  int64_t filter_launches;   /* step kernel launches inside t_filter_device */
  double t_eig;              /* host seconds in the dense eigensolves of the projected
                                operator (part of t_rr) */
} rcg_solve_result;

/* RayleighChebyshevLM::getMinEigenSystem (RayleighChebyshevLM.h:693-738;
 * RayleighChebyshevLMCuda.h:775-820) and getMinIntervalEigenSystem
 * (RayleighChebyshevLM.h:745-806).  evals and res (capacity `cap`) receive
 * the eigenvalues and residual norms; evecs (optional, cols >= cap) receives
 * the eigenvectors on the device; trace (optional, 10 doubles per inner
 * iteration: pass, iteration, degree, starBound, shift, lambdaStar,
 * maxResidual, maxEigDiff, eigDiffRatio, maxGap) up to trace_cap rows.
 * Returns the found count through result->found. */
int rcg_solve_lowest(rcg_operator op, const rcg_solve_params* params, rcg_solve_result* result,
                     double* evals, double* res, int64_t cap, rcg_block evecs, double* trace,
                     int64_t trace_cap);

/* ------------------------------------------------------------ serving */
/* No counterpart in the reference (it solves one operator per call): a
 * continuous-batching solve server on one device.  Requests queue in arrival
 * order and are admitted when one of `max_active` worker slots is free and
 * the request's device footprint fits `hbm_budget` bytes next to the running
 * solves and the resident eigenvectors of finished tickets (counted until
 * rcg_server_release); a request that does not fit runs once no other solve
 * is running.  Running solves share their Chebyshev filter launches (a batcher
 * gathers the sweeps of all running solves for up to `window_us`
 * microseconds); everything else runs on each worker's own stream.  A served
 * solve gives the same result as the same request solved alone.  The caller
 * keeps each operator alive until its ticket is done. */
int rcg_server_create(rcg_context ctx, int64_t hbm_budget, int32_t max_active, int32_t window_us,
                      rcg_server* out);
/* Queue a solve (params as rcg_solve_lowest; random start vectors). */
int rcg_server_submit(rcg_server s, rcg_operator op, const rcg_solve_params* params,
                      int64_t* ticket);
/* Warm start (setNonRandomStartFlag across requests, LM:957-985): the solve
 * starts from the resident eigenvectors of the finished ticket start_ticket
 * (same operator size and dtype), e.g. a nearby operator of a parameter
 * sweep.  start_ticket cannot be released until this solve has finished.
 * params->nonrandom_start must be 0 (the count comes from the start ticket). */
int rcg_server_submit_warm(rcg_server s, rcg_operator op, const rcg_solve_params* params,
                           int64_t start_ticket, int64_t* ticket);
/* Wait for a ticket: result, eigenvalues and residuals (capacity cap), and
 * times[3] = submit, admit, completion (seconds, steady clock).  Returns the
 * solve's own status. */
int rcg_server_wait(rcg_server s, int64_t ticket, rcg_solve_result* result, double* evals,
                    double* res, int64_t cap, double* times);
/* The ticket's eigenvectors, resident on the device (owned by the server
 * until rcg_server_release). */
int rcg_server_vectors(rcg_server s, int64_t ticket, rcg_block* evecs);
int rcg_server_release(rcg_server s, int64_t ticket);
/* stats[6]: completed, filter batches, requests in them, HBM in use, HBM
 * peak, queued. */
int rcg_server_stats(rcg_server s, int64_t* stats);
/* out[4] about the shared filter launches: device seconds of the finished
 * batches (first launch to last), kernel launches, algorithmic bytes of the
 * launched steps, batches timed. */
int rcg_server_filter_profile(rcg_server s, double* out);
/* Finishes the running solves, drops queued ones. */
int rcg_server_destroy(rcg_server s);

/* Sizes of the two structs (ABI check for bindings). */
int64_t rcg_sizeof_solve_params(void);
int64_t rcg_sizeof_solve_result(void);

#ifdef __cplusplus
}
#endif

#endif /* RCG_C_API_H_ */
