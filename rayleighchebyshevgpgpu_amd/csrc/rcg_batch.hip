// rcg_batch.hip -- the batched filter step of the serving loop: step j of
// the Chebyshev filters of many operators in one launch.  A translation unit
// of its own, so that its instantiations of csr_rows cannot change the code
// generated for the single-operator step in rcg_kernels.hip.
#include "rcg_csr_rows.hpp"

namespace rcg {
namespace {

static_assert(kBatchMaxRows <= kMaxRpb, "batched row blocks must fit the LDS row-pointer stage");

// Nonzeros per gather chunk of the batched step (the served operators' rows
// hold 5 or 7), measured separately: FP64 3 (32 served operators, degree 24:
// 8.39 ms per call; 2: 8.61, 4: 8.88), FP32 2 (the served stream's shared
// launches, 1024 requests: 8.67 s of device time; 3: 9.07-9.20 s, 4:
// 9.09-9.22 s).
#ifndef RCG_BATCH_CHUNK_F64
#define RCG_BATCH_CHUNK_F64 3
#endif
#ifndef RCG_BATCH_CHUNK_F32
#define RCG_BATCH_CHUNK_F32 2
#endif
template <typename T>
constexpr int batch_chunk() {
  return std::is_same<T, float>::value ? RCG_BATCH_CHUNK_F32 : RCG_BATCH_CHUNK_F64;
}

// ---------------------------------------------------- batched filter step
// Step j of the Chebyshev filters of many operators at once: task t is row
// block tasks[t].blk of request tasks[t].req.  Requests hold their own CSR
// arrays, blocks and coefficients; a request's blocks buf[0], buf[1] are the
// recurrence's ping-pong pair (X_{j-1} in buf[(j-1)%2], X_{j-2} and X_j in
// buf[j%2]; the first step reads buf[0] and writes buf[1]).  SRC: the FP32
// correction steps (j >= 2 only) with the source block R and the kap table
// (row j at kap + j*ld).  Workgroups stride over the tasks.
template <typename T, int G, int NV, bool SRC, bool CX>
__global__ __launch_bounds__(kThreads) void csr_batch_kernel(const BatchDesc* __restrict__ descs,
                                                             const BatchTask* __restrict__ tasks,
                                                             int ntasks, int j) {
  __shared__ int s_rp[kMaxRpb + 1];
  __shared__ int s_col[kStageNz];
  __shared__ T s_val[kStageNz * (CX ? 2 : 1)];
  const MixedArgs kNoStart{};  // the batched steps are never the mixed filter's start
  for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
    const BatchTask tk = tasks[t];
    const BatchDesc& d = descs[tk.req];
    const int64_t i0 = (int64_t)tk.blk * d.rows_per_blk;
    const int64_t i1 = min(d.n, i0 + d.rows_per_blk);
    const T* val = static_cast<const T*>(d.val);
    T* b0 = static_cast<T*>(d.buf[0]);
    T* b1 = static_cast<T*>(d.buf[1]);
    const T g1 = T(d.g1), g2 = T(d.g2);
    if constexpr (SRC) {
      T* xin = (j % 2 == 0) ? b1 : b0;
      T* xz = (j % 2 == 0) ? b0 : b1;
      csr_rows<T, G, NV, kChebSrc, 1, CX, batch_chunk<T>()>(d.n, d.rowptr, d.col, val, xin, xz, xz, d.ld, d.ncv, d.tail, g1,
                                          g2, T(1), i0, i1, static_cast<const T*>(d.R),
                                          static_cast<const T*>(d.kap) + (int64_t)j * d.ld, s_rp,
                                          s_col, s_val, kNoStart);
    } else if (j == 1) {
      csr_rows<T, G, NV, kChebFirst, 1, CX, batch_chunk<T>()>(d.n, d.rowptr, d.col, val, b0, nullptr, b1, d.ld,
                                            d.ncv, d.tail, g1, g2, T(d.m == 1 ? 0.5 : 1.0), i0, i1,
                                            nullptr, nullptr, s_rp, s_col, s_val, kNoStart);
    } else {
      T* xin = (j % 2 == 0) ? b1 : b0;
      T* xz = (j % 2 == 0) ? b0 : b1;
      const T sc = (j == d.m) ? T(1.0 / double(d.m + 1)) : T(1);
      csr_rows<T, G, NV, kCheb, 1, CX, batch_chunk<T>()>(d.n, d.rowptr, d.col, val, xin, xz, xz, d.ld, d.ncv, d.tail, g1, g2,
                                       sc, i0, i1, nullptr, nullptr, s_rp, s_col, s_val, kNoStart);
    }
    __syncthreads();  // the LDS stages are reused by the next task
  }
}

}  // namespace

// Launch classes of the batched step: lanes per row G = the row's 16-byte
// vectors up to 64, one vector per lane; G need not divide the workgroup
// (3, 6, 12, 24, 48), so no lane of a row sits idle at the served widths
// (k + max(4, k/2) columns, k in {8, 16, 32}: 3, 6 and 12 FP32 vectors, 6,
// 12 and 24 FP64).  Against power-of-two G with a quarter of the lanes idle:
// 32 served operators filtered together (FP64, degree 24) 10.1 -> 8.9 ms,
// the served stream 1203 -> 1254 eigenpairs/s.  Two vectors per lane on
// half the lanes (3x2, 6x2, 12x2, ...) lose again: 8.8 -> 9.6 ms, 1245 ->
// 1172 eigenpairs/s.
namespace {
constexpr int kBatchLim[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 128, 256};
constexpr int kBatchG[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 64, 64};
}  // namespace
// the launch switch below instantiates exactly these (G, vectors per lane)
static_assert(sizeof(kBatchLim) == sizeof(kBatchG), "one G per class");

template <typename T, bool SRC, bool CX>
void cheb_batch_t(const BatchDesc* descs, const BatchTask* tasks, int ntasks, int j, int cls,
                  hipStream_t st) {
  // one workgroup per task up to 8192 (workgroups striding over 1024 or 2048
  // tasks' worth served 1540 and 1580 eigenpairs/s against 1614)
  const unsigned grid = (unsigned)std::min(ntasks, 8192);
#define RCG_BATCH_CASE(C, GG, NN)                                                                 \
  case C:                                                                                         \
    hipLaunchKernelGGL((csr_batch_kernel<T, GG, NN, SRC, CX>), dim3(grid), dim3(kThreads), 0, st, \
                       descs, tasks, ntasks, j);                                                  \
    break;
  switch (cls) {
    RCG_BATCH_CASE(0, 1, 1)
    RCG_BATCH_CASE(1, 2, 1)
    RCG_BATCH_CASE(2, 3, 1)
    RCG_BATCH_CASE(3, 4, 1)
    RCG_BATCH_CASE(4, 6, 1)
    RCG_BATCH_CASE(5, 8, 1)
    RCG_BATCH_CASE(6, 12, 1)
    RCG_BATCH_CASE(7, 16, 1)
    RCG_BATCH_CASE(8, 24, 1)
    RCG_BATCH_CASE(9, 32, 1)
    RCG_BATCH_CASE(10, 48, 1)
    RCG_BATCH_CASE(11, 64, 1)
    RCG_BATCH_CASE(12, 64, 2)
    RCG_BATCH_CASE(13, 64, 4)
    default: throw Error(1, "bad CSR launch class");
  }
#undef RCG_BATCH_CASE
  RCG_HIP(hipGetLastError());
}

int batch_class(int ncv) {
  constexpr int nc = (int)(sizeof(kBatchLim) / sizeof(kBatchLim[0]));
  for (int c = 0; c < nc; ++c)
    if (ncv <= kBatchLim[c]) return c;
  throw Error(6, "CSR block kernels support at most 256 16-byte column vectors per row");
}

// Rows per task: the served stream's operators (n ~ 1e4) spread over twice
// the workgroups of 128-row tasks (served 1186 -> 1236-1258 eigenpairs/s;
// 32-row tasks 1213-1229).  The shared launches of 8 and 32 served requests
// alone (rcg_cheb_filter_batch, FP64, degree 24) run within 7% of each other
// for 32, 64 and 128 rows.  With the exact lane counts below, the served stream
// took 32, 64, 128 and 256 rows at 1574, 1580, 1543 and 1451 eigenpairs/s.  When G does not divide the workgroup, whole
// passes of kThreads / G rows nearest 64.
constexpr int kBatchRows = 64;
int batch_rows(int cls) {
  constexpr int nc = (int)(sizeof(kBatchLim) / sizeof(kBatchLim[0]));
  if (cls < 0 || cls >= nc) throw Error(1, "bad CSR launch class");
  const int g = kBatchG[cls];
  if (kThreads % g == 0) return kBatchRows;
  const int rs = kThreads / g;
  return rs * std::max(1, (kBatchRows + rs / 2) / rs);
}

void launch_cheb_batch(const BatchDesc* descs, const BatchTask* tasks, int ntasks, int j, bool f32,
                       bool src, bool cx, int cls, hipStream_t st) {
  if (ntasks <= 0) return;
  if (f32) {
    if (!src) throw Error(1, "FP32 batched steps are the correction steps");
    if (cx) cheb_batch_t<float, true, true>(descs, tasks, ntasks, j, cls, st);
    else cheb_batch_t<float, true, false>(descs, tasks, ntasks, j, cls, st);
  } else {
    if (src) throw Error(1, "FP64 batched steps are the plain recurrence");
    if (cx) cheb_batch_t<double, false, true>(descs, tasks, ntasks, j, cls, st);
    else cheb_batch_t<double, false, false>(descs, tasks, ntasks, j, cls, st);
  }
}

}  // namespace rcg
