// rcg_kernels.hip -- CDNA4 (gfx950, wave64) kernels of the Rayleigh-Chebyshev solve.
//
// Replaces LanczosCpolyOperatorLMCuda::apply (CPO-LMC:194-265: copy + SpMM +
// subtract per degree), the cuda_kernels.cu ops (scale, substract, normalize,
// residuals; CK:9-382), the cuBLAS projections (LMC:2139-2161, 2200-2222) and
// cuSOLVER QR (LMC:1979-2107) with kernels written for this chip:
//  * csr_block_kernel: one fused three-term Chebyshev step per degree.  A row
//    of the block (ld elements, 16-byte vectors) is owned by G lanes, so each
//    CSR gather is a single contiguous read of that row; launches are remapped
//    so that each XCD's L2 streams a contiguous range of rows.
//  * gram_partial_kernel / tall_gemm_kernel: LDS-tiled tall-skinny products
//    with 4x4 register tiles per thread; partial sums are reduced in a fixed
//    order so results do not depend on scheduling.
//  * column reductions (norms, residuals) and the Lanczos vector kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "rcg_csr_rows.hpp"

namespace rcg {
namespace {

template <typename T, int G, int NV, int MODE, int RPI, bool CX>
__global__ __launch_bounds__(kThreads) void csr_block_kernel(
    int64_t n, const int* __restrict__ rowptr, const int* __restrict__ col,
    const T* __restrict__ val, const T* __restrict__ X, const T* Z, T* out, int ld, int ncv,
    int tail, T p0, T p1, T p2, int64_t rows_per_blk, int64_t nblk, const int* __restrict__ order,
    int remap, const T* __restrict__ R, const T* __restrict__ kap, MixedArgs ms) {
  __shared__ int s_rp[kMaxRpb + 1];
  __shared__ int s_col[kStageNz];
  __shared__ T s_val[kStageNz * (CX ? 2 : 1)];
  const int64_t lq = remap ? xcd_remap(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
  if (lq >= nblk) return;  // uniform across the workgroup
  const int64_t lb = order ? (int64_t)order[lq] : lq;
  const int64_t i0 = lb * rows_per_blk;
  const int64_t i1 = min(n, i0 + rows_per_blk);
  csr_rows<T, G, NV, MODE, RPI, CX, csr_chunk<T, CX>()>(n, rowptr, col, val, X, Z, out, ld, ncv,
                                                       tail, p0, p1, p2, i0, i1, R, kap, s_rp,
                                                       s_col, s_val, ms);
}

#ifndef RCG_CSR_RPI
#define RCG_CSR_RPI 1
#endif
// lanes per row and vectors per lane for rows of 17-32 16-byte vectors
// (b = 64 FP64, b = 128 FP32)
#ifndef RCG_CSR32_G
#define RCG_CSR32_G 16
#endif
#ifndef RCG_CSR32_NV
#define RCG_CSR32_NV 2
#endif
// and for rows of 9-16 vectors (b = 64 FP32: 8 x 2 with the tuned 16x16 tile
// order 548 us per correction step at 128^3; 16 x 1 655 us, 4 x 4 718 us)
#ifndef RCG_CSR16_G
#define RCG_CSR16_G 8
#endif
#ifndef RCG_CSR16_NV
#define RCG_CSR16_NV 2
#endif


template <typename T, int MODE, bool CX, int G, int NV>
void csr_launch_g(int64_t n, const int* rp, const int* ci, const T* va, const T* X, const T* Z,
                  T* out, int ld, int ncv, int tail, T p0, T p1, T p2, hipStream_t st,
                  const CsrSched* sc, const T* R, const T* kap, const MixedArgs& ms) {
  constexpr int RS = kThreads / G;
  int64_t rows_per_blk, nblk;
  const int* order = nullptr;
  if (sc && sc->order) {
    rows_per_blk = sc->rows_per_blk;
    nblk = sc->nblk;
    order = sc->order;
  } else {
    // ~128-row blocks (their nonzeros fit the LDS stage for stencils) keep
    // the grid >> 256 CUs at large n; small problems use one row pass.
    rows_per_blk = std::min<int64_t>(std::max<int64_t>(RS, 128), kMaxRpb);
    nblk = (n + rows_per_blk - 1) / rows_per_blk;
    if (nblk < 2048) {
      rows_per_blk = RS;
      nblk = (n + rows_per_blk - 1) / rows_per_blk;
    }
  }
  if (rows_per_blk > kMaxRpb) throw Error(1, "CSR row block exceeds the staged row-pointer range");
  const int remap = sc ? sc->remap : 1;
  const int64_t grid = round_up(nblk, kXcds);
  hipLaunchKernelGGL((csr_block_kernel<T, G, NV, MODE, RCG_CSR_RPI, CX>), dim3((unsigned)grid), dim3(kThreads), 0,
                     st, n, rp, ci, va, X, Z, out, ld, ncv, tail, p0, p1, p2, rows_per_blk, nblk, order,
                     remap, R, kap, ms);
  RCG_HIP(hipGetLastError());
}

template <typename T, int MODE, bool CX>
void csr_launch_cx(int64_t n, const int* rp, const int* ci, const T* va, const T* X, const T* Z,
                   T* out, int ld, int ncols, T p0, T p1, T p2, hipStream_t st, const CsrSched* sc,
                   const T* R, const T* kap, const MixedArgs& ms) {
  constexpr int W = V16<T>::W;
  const int ncv = (ncols + W - 1) / W;
  const int tail = ncols - (ncv - 1) * W;
  if (ld % 4 != 0) throw Error(1, "block leading dimension must be a multiple of 4");
#define RCG_CSR_CASE(LIM, GG, NN) \
  if (ncv <= LIM) return csr_launch_g<T, MODE, CX, GG, NN>(n, rp, ci, va, X, Z, out, ld, ncv, tail, p0, p1, p2, st, sc, R, kap, ms);
  // (16-byte vectors per row, lanes per row G, vectors per lane NV).  From 8
  // vectors on, each lane carries two or more vectors: half the lanes per row
  // of one vector each, twice the rows per wave (FP64 step at 128^3, b = 64:
  // 942 -> 867 us; 1024^2 complex, b = 96: 1351 -> 1343 us).  Four vectors
  // per lane from 33 vectors on: the complex64 step at b = 96 (48 vectors),
  // 925 -> 709 us; four per lane at 16-32 vectors and eight at 65-128 were
  // slower.  A column's sum runs over the row's nonzeros in the same
  // order for any (G, NV), so results do not depend on this table.  The
  // batched step (rcg_batch.hip) keeps one vector per lane: the same change
  // there served 1168-1182 against 1181-1186 eigenpairs/s (and, while both
  // kernels shared this file, made the single-operator FP64 step 3% slower).
  // Lane counts that do not divide the workgroup (3, 6, 12, 24 lanes at 3-6,
  // 9-12, 17-24, 65-96 vectors and so on), which the batched step uses, were
  // flat or slower here: 1024^2 complex FP64 solve 29.7 -> 29.7, complex64
  // mixed 33.5 -> 32.3 eigenpairs/s, plain FP64 solves of served requests +1.4%.
  RCG_CSR_CASE(1, 1, 1)
  RCG_CSR_CASE(2, 2, 1)
  RCG_CSR_CASE(4, 4, 1)
  RCG_CSR_CASE(8, 4, 2)
  RCG_CSR_CASE(16, RCG_CSR16_G, RCG_CSR16_NV)
  RCG_CSR_CASE(32, RCG_CSR32_G, RCG_CSR32_NV)
  RCG_CSR_CASE(64, 16, 4)
  RCG_CSR_CASE(128, 32, 4)
  RCG_CSR_CASE(256, 32, 8)
#undef RCG_CSR_CASE
  throw Error(6, "CSR block kernels support at most 256 16-byte column vectors per row");
}

// ncols counts real columns (2 per complex column when cx).
template <typename T, int MODE>
void csr_launch(int64_t n, const int* rp, const int* ci, const T* va, const T* X, const T* Z,
                T* out, int ld, int ncols, T p0, T p1, T p2, hipStream_t st, const CsrSched* sc,
                const T* R = nullptr, const T* kap = nullptr, bool cx = false,
                const MixedArgs& ms = MixedArgs{}) {
  if (cx) {
    if (ncols % 2 != 0) throw Error(1, "complex blocks have an even number of real columns");
    csr_launch_cx<T, MODE, true>(n, rp, ci, va, X, Z, out, ld, ncols, p0, p1, p2, st, sc, R, kap, ms);
  } else {
    csr_launch_cx<T, MODE, false>(n, rp, ci, va, X, Z, out, ld, ncols, p0, p1, p2, st, sc, R, kap, ms);
  }
}

// ------------------------------------------------------ tall-skinny Gram
// Partial X^H Y over a row range for one 64x64 output tile.
constexpr int kTile = 64;
constexpr int kRb = 32;  // rows staged per LDS round

template <typename T>
__global__ __launch_bounds__(kThreads) void gram_partial_kernel(
    int64_t n, const T* __restrict__ X, int ldx, int bx, const T* __restrict__ Y, int ldy, int by,
    T* __restrict__ P, int64_t rows_per_blk, int ntb) {
  __shared__ T xs[kRb][kTile + 1];
  __shared__ T ys[kRb][kTile + 1];
  const int tile = blockIdx.y;
  const int ta = tile % ((bx + kTile - 1) / kTile);
  const int tb = tile / ((bx + kTile - 1) / kTile);
  const int a0 = ta * kTile;
  const int b0 = tb * kTile;
  const int tx = threadIdx.x % 16;
  const int ty = threadIdx.x / 16;
  const int64_t blk = blockIdx.x;
  const int64_t i0 = blk * rows_per_blk;
  const int64_t i1 = min(n, i0 + rows_per_blk);
  T acc[4][4];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[p][q] = T(0);
  for (int64_t r0 = i0; r0 < i1; r0 += kRb) {
#pragma unroll
    for (int e = threadIdx.x; e < kRb * kTile; e += kThreads) {
      const int rr = e / kTile;
      const int cc = e % kTile;
      const int64_t i = r0 + rr;
      const bool rowok = i < i1;
      xs[rr][cc] = (rowok && a0 + cc < bx) ? X[i * ldx + a0 + cc] : T(0);
      ys[rr][cc] = (rowok && b0 + cc < by) ? Y[i * ldy + b0 + cc] : T(0);
    }
    __syncthreads();
#pragma unroll 8
    for (int rr = 0; rr < kRb; ++rr) {
      T xa[4], yb[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) xa[p] = xs[rr][tx * 4 + p];
#pragma unroll
      for (int q = 0; q < 4; ++q) yb[q] = ys[rr][ty * 4 + q];
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] = fma(xa[p], yb[q], acc[p][q]);
    }
    __syncthreads();
  }
  T* Pb = P + blk * (int64_t)bx * by;
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int a = a0 + tx * 4 + p;
      const int b = b0 + ty * 4 + q;
      if (a < bx && b < by) Pb[a + (int64_t)b * bx] = acc[p][q];
    }
}

// out[o] = sum_b P[b*m + o].  A workgroup owns 64 outputs; its 4 waves take
// the partial blocks b = s (mod 4) with four interleaved accumulators each and
// are combined in a fixed order, so the result does not depend on scheduling.
template <typename T, typename TO>
__global__ __launch_bounds__(kThreads) void reduce_partials_kernel(const T* __restrict__ P,
                                                                   int64_t nblk, int64_t m,
                                                                   TO* __restrict__ out) {
  __shared__ T red[4][64];
  const int ol = threadIdx.x % 64;
  const int sl = threadIdx.x / 64;
  const int64_t o = blockIdx.x * (int64_t)64 + ol;
  T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
  if (o < m) {
    int64_t b = sl;
    // four rounds' loads in flight, added in the order of the one-round loop
    for (; b + 60 < nblk; b += 64) {
      T x[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) x[k] = P[(b + 4 * k) * m + o];
#pragma unroll
      for (int k = 0; k < 16; k += 4) {
        a0 += x[k];
        a1 += x[k + 1];
        a2 += x[k + 2];
        a3 += x[k + 3];
      }
    }
    for (; b + 12 < nblk; b += 16) {
      a0 += P[b * m + o];
      a1 += P[(b + 4) * m + o];
      a2 += P[(b + 8) * m + o];
      a3 += P[(b + 12) * m + o];
    }
    for (; b < nblk; b += 4) a0 += P[b * m + o];
  }
  red[sl][ol] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  if (sl == 0 && o < m) out[o] = TO((red[0][ol] + red[1][ol]) + (red[2][ol] + red[3][ol]));
}

inline unsigned reduce_grid(int64_t m) { return (unsigned)((m + 63) / 64); }

template <typename T>
int64_t gram_blocks(int64_t n, int bx, int by, int64_t* rows_per_blk) {
  const int ntiles = ((bx + kTile - 1) / kTile) * ((by + kTile - 1) / kTile);
  // row blocks: 1024 over the tiles, 512 from 2^20 rows (half the partials to
  // write and reduce: the 128^3 x 64 FP64 solve 1977 -> 1967 ms; 256: 1973 ms)
  const int64_t blocks = n >= (int64_t(1) << 20) ? 512 : 1024;
  int64_t target = std::max<int64_t>(1, blocks / ntiles);
  int64_t rpb = round_up((n + target - 1) / target, kRb);
  // at least 4 rows per output column, so a block's bx x by partial is at
  // most a quarter of the rows it reads (small n would otherwise write about
  // as many partial bytes as it reads), and at least 256 rows: the small
  // operators of a served stream (n ~ 1e4) then launch tens of workgroups with
  // fewer partials to reduce (served throughput +4%; 512 and 1024 were slower)
  rpb = std::max<int64_t>(rpb, round_up(4 * (int64_t)std::max(bx, by), kRb));
  rpb = std::max<int64_t>(rpb, 256);
  *rows_per_blk = rpb;
  return (n + rpb - 1) / rpb;
}

// -------------------------------------------------------- tall x small GEMM
template <typename T>
__global__ __launch_bounds__(kThreads) void tall_gemm_kernel(
    int64_t n, const T* __restrict__ X, int ldx, int bx, const T* __restrict__ C, int ldc, int by,
    T alpha, T beta, const T* Z, T* out, int ldo, int64_t nrb) {
  constexpr int KC = 16;
  __shared__ T xs[kTile][KC + 1];
  __shared__ T cs[KC][kTile + 1];
  const int64_t rb = xcd_remap(blockIdx.x, gridDim.x);
  if (rb >= nrb) return;
  const int c0 = blockIdx.y * kTile;
  const int tc = threadIdx.x % 16;
  const int tr = threadIdx.x / 16;
  const int64_t i0 = rb * kTile;
  T acc[4][4];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[p][q] = T(0);
  for (int k0 = 0; k0 < bx; k0 += KC) {
    for (int e = threadIdx.x; e < kTile * KC; e += kThreads) {
      const int rr = e / KC;
      const int kk = e % KC;
      const int64_t i = i0 + rr;
      xs[rr][kk] = (i < n && k0 + kk < bx) ? X[i * ldx + k0 + kk] : T(0);
    }
    for (int e = threadIdx.x; e < KC * kTile; e += kThreads) {
      const int kk = e / kTile;
      const int cc = e % kTile;
      cs[kk][cc] = (k0 + kk < bx && c0 + cc < by) ? C[(k0 + kk) + (int64_t)(c0 + cc) * ldc] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) {
      T xr[4], cr[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) xr[p] = xs[tr * 4 + p][kk];
#pragma unroll
      for (int q = 0; q < 4; ++q) cr[q] = cs[kk][tc * 4 + q];
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] = fma(xr[p], cr[q], acc[p][q]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int64_t i = i0 + tr * 4 + p;
    if (i >= n) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = c0 + tc * 4 + q;
      if (c >= by) continue;
      const int64_t off = i * ldo + c;
      T v = alpha * acc[p][q];
      if (beta != T(0)) v = fma(beta, Z[off], v);
      out[off] = v;
    }
  }
}

// ---------------------------------------------------- FP64 MFMA versions
// v_mfma_f64_16x16x4f64 computes D(16x16) += A(16x4) B(4x16); lane l supplies
// A[l%16][l/16] and B[l/16][l%16] and holds D[4j + l/16][l%16] in register j.
typedef double d4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4v mfma16x16x4(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Which tile column lane cc of a 16x16 subtile t holds.  With an even number
// of subtiles the columns are taken in pairs: lane cc reads columns 2cc and
// 2cc+1 of each 32-column half with one 16-byte load, subtile 2h holding the
// even and 2h+1 the odd ones; otherwise subtile t holds columns 16t..16t+15.
// The output tile is the same sums with its rows and columns permuted.
template <int NT>
__device__ __forceinline__ int gram_col(int t, int cc) {
  if constexpr (NT % 2 == 0) return 32 * (t >> 1) + 2 * cc + (t & 1);
  return 16 * t + cc;
}
// the subtile that holds tile column c (the inverse of gram_col)
template <int NT>
__device__ __forceinline__ int gram_sub(int c) {
  if constexpr (NT % 2 == 0) return 2 * (c >> 5) + (c & 1);
  return c >> 4;
}

// Partial X^T Y of one 64x64 output tile over a row range.  Wave w takes the
// 4-row groups i0 + 4w + 16t; lane (kr = l/16, cc = l%16) loads row row+kr of
// X at the tile columns gram_col(p, cc) and of Y at gram_col(q, cc), which are
// the A and B fragments of the NT x NT MFMA subtiles (16-byte loads of column
// pairs when NT is even, 128-byte row segments across the 16 lanes either
// way).  A wave's rows and the order it sums them in do not depend on the
// column order, so every entry is the same sum as with 8-byte loads.  DIAG (a
// tile on the diagonal of X^T X) multiplies only subtiles p <= q and needs
// no Y loads.
template <bool DIAG, int NT>
__device__ __forceinline__ void gram_tile_accumulate(d4v (&acc)[NT][NT], const double* __restrict__ X,
                                                     int ldx, int bx, const double* __restrict__ Y,
                                                     int ldy, int by, int a0, int b0, int64_t ib,
                                                     int64_t i1, int kr, int cc) {
  // U 16-row groups per pass: all of their loads are issued before the MFMAs
  // that consume them, and the next pass's loads before this pass's MFMAs.
  // A wave's rows and the order it sums them in do not depend on U.  Two
  // groups at NT = 3 (142 instead of 198 VGPRs for X^T Y: served stream
  // 1602 -> 1611 eigenpairs/s); at NT = 4 two groups made the 128^3 x 64
  // solve 1975 -> 1989 ms.
  constexpr int U = NT == 3 ? 2 : 4;
  double xa[U][NT], yb[U][NT], xn[U][NT], yn[U][NT];
  // column pairs are read with one 16-byte load where both columns exist and
  // the rows are 16-byte aligned (a0, b0 are multiples of 64)
  const bool vx = NT % 2 == 0 && ((reinterpret_cast<uintptr_t>(X) & 15) == 0) && (ldx % 2 == 0);
  const bool vy = NT % 2 == 0 && ((reinterpret_cast<uintptr_t>(Y) & 15) == 0) && (ldy % 2 == 0);
  auto row_load = [&](const double* __restrict__ B, int ldb, int bw, int c0, bool vec, int64_t row,
                      bool ok, double (&v)[NT]) {
    if constexpr (NT % 2 == 0) {
#pragma unroll
      for (int h = 0; h < NT / 2; ++h) {
        const int c = c0 + gram_col<NT>(2 * h, cc);
        if (ok && vec && c + 1 < bw) {
          const dvec2 t = *reinterpret_cast<const dvec2*>(B + row * ldb + c);
          v[2 * h] = t.x;
          v[2 * h + 1] = t.y;
        } else {
          v[2 * h] = (ok && c < bw) ? B[row * ldb + c] : 0.0;
          v[2 * h + 1] = (ok && c + 1 < bw) ? B[row * ldb + c + 1] : 0.0;
        }
      }
    } else {
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int c = c0 + gram_col<NT>(t, cc);
        v[t] = (ok && c < bw) ? B[row * ldb + c] : 0.0;
      }
    }
  };
  auto load = [&](int64_t jb, double (&xv)[U][NT], double (&yv)[U][NT]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t row = jb + 16 * u + kr;
      const bool ok = row < i1;
      row_load(X, ldx, bx, a0, vx, row, ok, xv[u]);
      if constexpr (!DIAG) row_load(Y, ldy, by, b0, vy, row, ok, yv[u]);
    }
  };
  if (ib < i1) load(ib, xa, yb);
  for (; ib < i1; ib += 16 * U) {
    const bool more = ib + 16 * U < i1;
    if (more) load(ib + 16 * U, xn, yn);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int p = 0; p < NT; ++p)
#pragma unroll
        for (int q = 0; q < NT; ++q) {
          if (DIAG && p > q) continue;
          acc[p][q] = mfma16x16x4(xa[u][p], DIAG ? xa[u][q] : yb[u][q], acc[p][q]);
        }
    if (more) {
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          xa[u][t] = xn[u][t];
          if constexpr (!DIAG) yb[u][t] = yn[u][t];
        }
    }
  }
}

// Which 64x64 tiles of X^H Y a Gram launch multiplies.  kGramFull: all of
// them.  kGramSym (X == Y): only the tiles ta <= tb of the upper triangle,
// ntiles = nta (nta+1)/2, each written at (a, b) and mirrored at (b, a).
// kGramLower (bx == by, X != Y): the tiles on and below the diagonal, for
// callers that read only the lower triangle of X^H Y (the Hermitian
// projection); the upper tiles get the transposed lower values so that every
// entry of the result is defined.  kGramDiag: the one tile of X^T X when
// bx <= 64.  kGramPair (bx == by <= 64): X^T X and X^T Y in one launch, tile 0
// and tile 1 of each row block, with the partials of both side by side.
enum { kGramFull = 0, kGramSym = 1, kGramDiag = 2, kGramLower = 3, kGramPair = 4 };

// Partial Gram matrices of one 64x64 tile, one bx x by block per row block:
// the one-tile launches (kGramFull with bx, by <= 64, or kGramDiag, compiled
// without the general path so that it keeps two waves per SIMD; kGramPair
// runs either per workgroup).  A row block's partials land at
// P + (rblk * ntiles + tile) * bx * by.  NT = ceil(max(bx, by) / 16): only
// the NT x NT 16x16 subtiles that hold output are multiplied, with registers
// for those alone (the sums of the others were zero).
template <int MODE, int NT>
__global__ __launch_bounds__(kThreads) void gram_mfma_kernel(
    int64_t n, const double* __restrict__ X, int ldx, int bx, const double* __restrict__ Y, int ldy,
    int by, double* __restrict__ P, int64_t rows_per_blk, int nta, int ntiles, int64_t nblk) {
  static_assert(MODE == kGramFull || MODE == kGramDiag || MODE == kGramPair,
                "one-tile launches only");
  __shared__ double red[kTile * kTile];
  // one 1-D grid: the tiles of a row block are consecutive logical workgroups,
  // which the remap keeps on one XCD, so its rows are read from HBM once
  const int64_t L = xcd_remap(blockIdx.x, gridDim.x);
  if (L >= nblk * ntiles) return;
  const int tile = (int)(L % ntiles);
  const int64_t rblk = L / ntiles;
  const int ta = MODE == kGramPair ? 0 : tile % nta;
  const int tb = MODE == kGramPair ? 0 : tile / nta;
  // uniform per workgroup: kGramPair's tile 0 is X^T X
  const bool diag = MODE == kGramDiag || (MODE == kGramPair && tile == 0);
  const int a0 = ta * kTile;
  const int b0 = tb * kTile;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kr = lane >> 4;
  const int cc = lane & 15;
  const int64_t i0 = rblk * rows_per_blk;
  const int64_t i1 = min(n, i0 + rows_per_blk);
  d4v acc[NT][NT];
#pragma unroll
  for (int p = 0; p < NT; ++p)
#pragma unroll
    for (int q = 0; q < NT; ++q) acc[p][q] = d4v{0.0, 0.0, 0.0, 0.0};
  if constexpr (MODE == kGramPair) {
    if (diag)
      gram_tile_accumulate<true, NT>(acc, X, ldx, bx, X, ldx, bx, a0, b0, i0 + 4 * wave, i1, kr,
                                     cc);
    else
      gram_tile_accumulate<false, NT>(acc, X, ldx, bx, Y, ldy, by, a0, b0, i0 + 4 * wave, i1, kr,
                                      cc);
  } else {
    gram_tile_accumulate<MODE == kGramDiag, NT>(acc, X, ldx, bx, Y, ldy, by, a0, b0,
                                                i0 + 4 * wave, i1, kr, cc);
  }
  // the four waves' tiles are summed in LDS in wave order
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int p = 0; p < NT; ++p)
#pragma unroll
        for (int q = 0; q < NT; ++q) {
          if (diag && p > q) continue;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int idx = gram_col<NT>(p, 4 * j + kr) + kTile * gram_col<NT>(q, cc);
            red[idx] = (w == 0) ? acc[p][q][j] : red[idx] + acc[p][q][j];
          }
        }
    }
    __syncthreads();
  }
  double* Pb = P + (rblk * (MODE == kGramPair ? ntiles : 1) + (MODE == kGramPair ? tile : 0)) *
                      (int64_t)bx * by;
  for (int e = threadIdx.x; e < kTile * kTile; e += kThreads) {
    const int al = e % kTile;
    const int bl = e / kTile;
    const int a = a0 + al;
    const int b = b0 + bl;
    if (a >= bx || b >= by) continue;
    // a diagonal tile holds the subtiles p <= q; the others are their mirror
    if (diag)
      Pb[a + (int64_t)b * bx] = (gram_sub<NT>(al) <= gram_sub<NT>(bl)) ? red[al + kTile * bl]
                                                                        : red[bl + kTile * al];
    else
      Pb[a + (int64_t)b * bx] = red[al + kTile * bl];
  }
}

// Gram partials for blocks wider than one tile (bx or by > 64), through LDS.
// Each workgroup owns one 64x64 output tile (enumerated as in gram_mfma_kernel)
// and one row block; 32-row chunks of X[:, a0:a0+64] and Y[:, b0:b0+64] are
// loaded with 16-byte loads one chunk ahead and staged in double-buffered LDS
// (one barrier per chunk; loads two chunks ahead, as in the dense step, were
// 5% faster for X^T X and 3-5% slower for the other two forms at 1024^2 x 192).  Wave w owns the 32x32 quadrant (w & 1, w >> 1) of
// the tile, so no cross-wave reduction is needed; in a diagonal tile of X^T X
// the quadrant below the diagonal is the mirror of the one above and its wave
// idles.  (Each wave multiplying a quarter of every chunk's rows into the
// whole tile, four MFMAs per fragment read instead of one and no idle wave,
// then a sum over the waves in LDS, measured the same at 1024^2 x 192.)  Row
// blocks are sized so that all workgroups are resident at once
// (two or four times as many smaller blocks, so that compute units done with
// the cheaper diagonal tiles take more, gave the same time at 1024^2 x 192: a
// workgroup's time is its chunks' latency, not its compute unit's MFMA load).
constexpr int kGramKr = 32;
template <int MODE>
__global__ __launch_bounds__(kThreads, 2) void gram_lds_kernel(
    int64_t n, const double* __restrict__ X, int ldx, int bx, const double* __restrict__ Y, int ldy,
    int by, double* __restrict__ P, int64_t rows_per_blk, int nta, int ntiles, int64_t nblk) {
  __shared__ double xs[2][kGramKr][kTile + 1];
  __shared__ double ys[2][kGramKr][kTile + 1];
  const int64_t L = xcd_remap(blockIdx.x, gridDim.x);
  if (L >= nblk * ntiles) return;
  const int tile = (int)(L % ntiles);
  const int64_t rblk = L / ntiles;
  int ta, tb;
  if constexpr (MODE != kGramFull) {
    tb = 0;
    while ((tb + 1) * (tb + 2) / 2 <= tile) ++tb;
    ta = tile - tb * (tb + 1) / 2;
    if constexpr (MODE == kGramLower) {
      const int t = ta;
      ta = tb;
      tb = t;
    }
  } else {
    ta = tile % nta;
    tb = tile / nta;
  }
  const bool diag = MODE == kGramSym && ta == tb;
  const int a0 = ta * kTile;
  const int b0 = tb * kTile;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kr = lane >> 4;
  const int cc = lane & 15;
  const int qa = wave & 1;
  const int qb = wave >> 1;
  const bool idle = diag && qa == 1 && qb == 0;
  const int64_t i0 = rblk * rows_per_blk;
  const int64_t i1 = min(n, i0 + rows_per_blk);
  const bool vx = ((reinterpret_cast<uintptr_t>(X) & 15) == 0) && (ldx % 2 == 0);
  const bool vy = ((reinterpret_cast<uintptr_t>(Y) & 15) == 0) && (ldy % 2 == 0);
  // thread t moves pairs e = t + 256 m, m < 4: chunk row e / 32, columns
  // 2 (e % 32) and 2 (e % 32) + 1 of the tile's 64
  auto ld2 = [&](const double* B, int ldb, int bw, bool vec, int64_t row, int col) -> dvec2 {
    if (row >= i1) return dvec2{0.0, 0.0};
    const double* ptr = B + row * ldb + col;
    if (vec && col + 1 < bw) return *reinterpret_cast<const dvec2*>(ptr);
    return dvec2{col < bw ? ptr[0] : 0.0, col + 1 < bw ? ptr[1] : 0.0};
  };
  dvec2 px[4], py[4];
  auto fetch = [&](int64_t r0) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int e = threadIdx.x + kThreads * m;
      const int64_t row = r0 + e / 32;
      const int c = 2 * (e % 32);
      px[m] = ld2(X, ldx, bx, vx, row, a0 + c);
      if (!diag) py[m] = ld2(Y, ldy, by, vy, row, b0 + c);
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int e = threadIdx.x + kThreads * m;
      const int rr = e / 32;
      const int c = 2 * (e % 32);
      xs[buf][rr][c] = px[m].x;
      xs[buf][rr][c + 1] = px[m].y;
      if (!diag) {
        ys[buf][rr][c] = py[m].x;
        ys[buf][rr][c + 1] = py[m].y;
      }
    }
  };
  d4v acc[2][2];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q) acc[p][q] = d4v{0.0, 0.0, 0.0, 0.0};
  const bool dquad = diag && qa == qb;  // quadrant on the diagonal: subtiles p <= q
  int buf = 0;
  if (i0 < i1) fetch(i0);
  for (int64_t r0 = i0; r0 < i1; r0 += kGramKr) {
    stash(buf);
    __syncthreads();
    if (r0 + kGramKr < i1) fetch(r0 + kGramKr);
    if (!idle) {
      const double(*yb)[kTile + 1] = diag ? xs[buf] : ys[buf];
#pragma unroll
      for (int k4 = 0; k4 < kGramKr; k4 += 4) {
        const int kk = k4 + kr;
        double af[2], bf[2];
#pragma unroll
        for (int p = 0; p < 2; ++p) af[p] = xs[buf][kk][32 * qa + 16 * p + cc];
#pragma unroll
        for (int q = 0; q < 2; ++q) bf[q] = yb[kk][32 * qb + 16 * q + cc];
        acc[0][0] = mfma16x16x4(af[0], bf[0], acc[0][0]);
        acc[0][1] = mfma16x16x4(af[0], bf[1], acc[0][1]);
        if (!dquad) acc[1][0] = mfma16x16x4(af[1], bf[0], acc[1][0]);
        acc[1][1] = mfma16x16x4(af[1], bf[1], acc[1][1]);
      }
    }
    buf ^= 1;
  }
  if (idle) return;
  double* Pb = P + rblk * (int64_t)bx * by;
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (dquad && p > q) continue;
      const int as = a0 + 32 * qa + 16 * p;
      const int bs = b0 + 32 * qb + 16 * q;
      const bool mirror = (MODE == kGramSym && as != bs) || (MODE == kGramLower && ta != tb);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int a = as + 4 * j + kr;
        const int b = bs + cc;
        if (a >= bx || b >= by) continue;
        const double v = acc[p][q][j];
        Pb[a + (int64_t)b * bx] = v;
        if (mirror) Pb[b + (int64_t)a * bx] = v;
      }
    }
}

// out = alpha X C + beta Z (X: n x bx, C: bx x by column-major).  A wave owns
// 16 rows of a 64-row tile; lane (kk = l/16, r = l%16) loads the 16
// consecutive entries [16kk, 16kk+16) of row r of a 64-wide k chunk with
// 16-byte loads, and MFMA step s contracts k = 16kk + s: A[r][kk] =
// X[r][16kk+s], B[kk][c] = C[16kk+s][c] (LDS).  D[4j+kk][c] lands in register
// j, so 16 lanes store 16 consecutive columns of one row.  Workgroups loop
// over row tiles; C is staged once when bx <= 64.  blockIdx.y selects 64
// output columns.
template <bool BETA, int NTC>
__global__ __launch_bounds__(kThreads) void tall_gemm_mfma_kernel(
    int64_t n, const double* __restrict__ X, int ldx, int bx, const double* __restrict__ C,
    int ldc, int by, double alpha, double beta, const double* Z, double* out, int ldo,
    int64_t ntiles, int ct, int64_t nstripe) {
  __shared__ double cs[kTile][kTile + 1];  // [k][c]
  // 1-D grid of nstripe x ct workgroups: the ct column tiles of a row stripe are
  // consecutive logical workgroups on one XCD, so X rows are read from HBM once
  const int64_t L = xcd_remap(blockIdx.x, gridDim.x);
  if (L >= nstripe * ct) return;
  const int c0 = (int)(L % ct) * kTile;
  const int64_t stripe = L / ct;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kk = lane >> 4;
  const int r = lane & 15;
  const int nk = (bx + kTile - 1) / kTile;
  const bool vec_ok = ((reinterpret_cast<uintptr_t>(X) & 15) == 0) && (ldx % 2 == 0);
  auto stage = [&](int k0) {
    for (int e = threadIdx.x; e < kTile * kTile; e += kThreads) {
      const int k = e % kTile;
      const int c = e / kTile;
      cs[k][c] = (k0 + k < bx && c0 + c < by) ? C[(k0 + k) + (int64_t)(c0 + c) * ldc] : 0.0;
    }
  };
  if (nk == 1) {
    stage(0);
    __syncthreads();
  }
  for (int64_t tile = stripe; tile < ntiles; tile += nstripe) {
    const int64_t row = tile * kTile + wave * 16 + r;
    d4v acc[NTC];
#pragma unroll
    for (int t = 0; t < NTC; ++t) acc[t] = d4v{0.0, 0.0, 0.0, 0.0};
    for (int ci = 0; ci < nk; ++ci) {
      const int k0 = ci * kTile;
      if (nk > 1) {
        __syncthreads();
        stage(k0);
        __syncthreads();
      }
      double xa[16];
      const int kb = k0 + 16 * kk;
      if (row < n && kb + 16 <= bx && vec_ok) {
        const double* xp = X + row * ldx + kb;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const dvec2 t = *reinterpret_cast<const dvec2*>(xp + 2 * q);
          xa[2 * q] = t.x;
          xa[2 * q + 1] = t.y;
        }
      } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) xa[q] = (row < n && kb + q < bx) ? X[row * ldx + kb + q] : 0.0;
      }
#pragma unroll
      for (int sidx = 0; sidx < 16; ++sidx) {
        // step sidx contracts k0 + sidx + {0, 16, 32, 48}: none is a column of X
        // once k0 + sidx >= bx (uniform)
        if (NTC == 4 || k0 + sidx < bx) {
          const int k = 16 * kk + sidx;
#pragma unroll
          for (int t = 0; t < NTC; ++t) acc[t] = mfma16x16x4(xa[sidx], cs[k][16 * t + r], acc[t]);
        }
      }
    }
    // Z may be `out` itself: all of the lane's Z values are read before its
    // first store, so the loads are issued together
    double zv[4][NTC];
    if constexpr (BETA) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t ro = tile * kTile + wave * 16 + 4 * j + kk;
#pragma unroll
        for (int t = 0; t < NTC; ++t) {
          const int c = c0 + 16 * t + r;
          zv[j][t] = (ro < n && c < by) ? Z[ro * ldo + c] : 0.0;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t ro = tile * kTile + wave * 16 + 4 * j + kk;
      if (ro >= n) continue;
#pragma unroll
      for (int t = 0; t < NTC; ++t) {
        const int c = c0 + 16 * t + r;
        if (c >= by) continue;
        double v = alpha * acc[t][j];
        if constexpr (BETA) v = fma(beta, zv[j][t], v);
        out[ro * ldo + c] = v;
      }
    }
  }
}

// out = alpha X C + beta Z for 64 < bx <= 64 KCH (the Gram-sized products of
// blocks of up to 256 columns).  The workgroup's whole 64-column slice of C
// (64 KCH x 64) is staged in LDS once; the workgroup then walks 128-row tiles
// with no further barrier.  Each wave owns 32 rows (two 16-row MFMA groups),
// so every B fragment read from LDS feeds two MFMAs, and the next 64-wide k
// chunk of its X rows -- in the last chunk, the next tile's first -- is loaded
// while the current one is multiplied.  Whether every row and column of a
// wave's chunk exists is decided once per wave (a scalar branch): loads that
// lanes take or skip on their own would make the wave wait for the loads in
// flight where the paths join, which takes the prefetch's overlap away.
template <int KCH>
__global__ __launch_bounds__(kThreads) void tall_gemm_wide_kernel(
    int64_t n, const double* __restrict__ X, int ldx, int bx, const double* __restrict__ C,
    int ldc, int by, double alpha, double beta, const double* Z, double* out, int ldo,
    int64_t ntiles, int ct, int64_t nstripe) {
  constexpr int KP = KCH * kTile;
  __shared__ double cs[KP][kTile + 1];  // [k][c]
  const int64_t L = xcd_remap(blockIdx.x, gridDim.x);
  if (L >= nstripe * ct) return;
  const int c0 = (int)(L % ct) * kTile;
  const int64_t stripe = L / ct;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kk = lane >> 4;
  const int r = lane & 15;
  const bool vec_ok = ((reinterpret_cast<uintptr_t>(X) & 15) == 0) && (ldx % 2 == 0);
  for (int e = threadIdx.x; e < KP * kTile; e += kThreads) {
    const int k = e % KP;
    const int c = e / KP;
    cs[k][c] = (k < bx && c0 + c < by) ? C[k + (int64_t)(c0 + c) * ldc] : 0.0;
  }
  __syncthreads();
  // lane (kk, r) of group g holds X[row_g][k0 + 16kk + q], q < 16; r0 and k0
  // are the same for the whole wave
  auto load = [&](int64_t r0, int k0, double (&xv)[2][16]) {
    const int kb = k0 + 16 * kk;
    const bool full = __builtin_amdgcn_readfirstlane(
                          (int)(vec_ok && r0 + 32 <= n && k0 + kTile <= bx)) != 0;
    if (full) {
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        const double* xp = X + (r0 + 16 * g + r) * ldx + kb;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const dvec2 t = *reinterpret_cast<const dvec2*>(xp + 2 * q);
          xv[g][2 * q] = t.x;
          xv[g][2 * q + 1] = t.y;
        }
      }
    } else {
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        const int64_t row = r0 + 16 * g + r;
#pragma unroll
        for (int q = 0; q < 16; ++q)
          xv[g][q] = (row < n && kb + q < bx) ? X[row * ldx + kb + q] : 0.0;
      }
    }
  };
  // xn receives every load and xa is only ever written from it by copies: a
  // register that entered the loops with a load in flight would make every
  // chunk wait for the loads issued ahead of its MFMAs
  double xa[2][16], xn[2][16];
  if (stripe < ntiles) load(stripe * (2 * kTile) + wave * 32, 0, xn);
  for (int64_t tile = stripe; tile < ntiles; tile += nstripe) {
    const int64_t r0 = tile * (2 * kTile) + wave * 32;
    const bool next = tile + nstripe < ntiles;
    // every output row and column of the wave's tile exists (Z reads and the
    // stores then take no per-lane branches)
    const bool ofull =
        __builtin_amdgcn_readfirstlane((int)(r0 + 32 <= n && c0 + kTile <= by)) != 0;
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int q = 0; q < 16; ++q) xa[g][q] = xn[g][q];
    d4v acc[2][4];
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[g][t] = d4v{0.0, 0.0, 0.0, 0.0};
    // beta Z of the epilogue, read at the start of the last chunk: Z may be
    // `out` itself, so loads issued between the stores would each wait for
    // the one before
    double zv[2][4][4];
#pragma unroll 1
    for (int ci = 0; ci < KCH; ++ci) {
      // one load site for both prefetches, so that they land in the same
      // registers and need no copies (which would wait for them) to merge
      if (ci + 1 < KCH || next)
        load(ci + 1 < KCH ? r0 : r0 + nstripe * (2 * kTile), ci + 1 < KCH ? (ci + 1) * kTile : 0,
             xn);
      if (ci + 1 == KCH && beta != 0.0) {
        if (ofull) {
#pragma unroll
          for (int g = 0; g < 2; ++g)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
              for (int t = 0; t < 4; ++t)
                zv[g][j][t] = Z[(r0 + 16 * g + 4 * j + kk) * ldo + c0 + 16 * t + r];
        } else {
#pragma unroll
          for (int g = 0; g < 2; ++g)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int64_t ro = r0 + 16 * g + 4 * j + kk;
#pragma unroll
              for (int t = 0; t < 4; ++t) {
                const int c = c0 + 16 * t + r;
                zv[g][j][t] = (ro < n && c < by) ? Z[ro * ldo + c] : 0.0;
              }
            }
        }
      }
#pragma unroll
      for (int sidx = 0; sidx < 16; ++sidx) {
        const int k = ci * kTile + 16 * kk + sidx;
        double b[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) b[t] = cs[k][16 * t + r];
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
          for (int t = 0; t < 4; ++t) acc[g][t] = mfma16x16x4(xa[g][sidx], b[t], acc[g][t]);
      }
      if (ci + 1 < KCH) {
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
          for (int q = 0; q < 16; ++q) xa[g][q] = xn[g][q];
      }
    }
    if (ofull) {
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            double v = alpha * acc[g][t][j];
            if (beta != 0.0) v = fma(beta, zv[g][j][t], v);
            out[(r0 + 16 * g + 4 * j + kk) * ldo + c0 + 16 * t + r] = v;
          }
    } else {
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int64_t ro = r0 + 16 * g + 4 * j + kk;
          if (ro >= n) continue;
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int c = c0 + 16 * t + r;
            if (c >= by) continue;
            double v = alpha * acc[g][t][j];
            if (beta != 0.0) v = fma(beta, zv[g][j][t], v);
            out[ro * ldo + c] = v;
          }
        }
    }
  }
}

// ------------------------------------------------ dense operator (MFMA)
// v_mfma_f32_16x16x4f32: lane l supplies A[l%16][l/16] and B[l/16][l%16] as
// for FP64; register j holds D[4(l/16) + j][l%16].
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f4v mfma16x16x4(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T>
struct MfmaAcc;
template <>
struct MfmaAcc<double> {
  typedef d4v type;
  // row within the 16x16 tile of register j of lane l
  __device__ static int row(int j, int l) { return 4 * j + (l >> 4); }
};
template <>
struct MfmaAcc<float> {
  typedef f4v type;
  __device__ static int row(int j, int l) { return 4 * (l >> 4) + j; }
};

// One filter step (or apply) of a dense symmetric operator A (n x n,
// row-major = column-major) on a row-interleaved block:
//   kApply: out = p0 A X;  kApplyBeta: out = p0 A X + p1 out;
//   kChebFirst: out = p2 p1 (p0 A X + X);  kCheb: out = p2 (p1 (p0 A X + X) - Z);
//   kChebSrc: kCheb + kap[c] R.
// BM x 64 output tile per workgroup; wave w owns rows [w BM/4, (w+1) BM/4), all
// 64 columns: FP64 (BM = 128) 2 x 4 MFMA 16x16x4 tiles, 6 LDS fragment reads
// per 8 MFMAs; FP32 (BM = 64) 1 x 4 tiles.
// K runs in chunks of BK through two LDS buffers and two register sets: a
// chunk's global loads are issued two chunks before its MFMAs and stored into
// the free LDS buffer after the MFMAs of the chunk before it, one barrier per
// chunk.  The column tiles of a row
// tile are consecutive logical workgroups, kept on one XCD, so the A rows
// they share come from the same L2.
template <typename T>
struct DenseCfg {
  // FP32 runs 64-row tiles (16 rows per wave): at n = 16384, b = 256 that is
  // 1024 workgroups instead of 512, 4.5% faster per step than 128 rows; 32-column
  // tiles, K chunks of 16 or 64, and 32x32 per wave were slower
  // (FP64 in 64-row tiles too, 4 waves per SIMD: 2.44 -> 2.69 ms per step)
  static constexpr int BM = sizeof(T) == 8 ? 128 : 64;
  static constexpr int BN = 64;
  static constexpr int BK = sizeof(T) == 8 ? 16 : 32;
  // LDS row padding (elements) of the A and X tiles.  FP32: rows of 34 and 80
  // floats put the 16 rows x 2 K of a half-wave's A fragment reads, and the
  // 16 columns x 2 K of its X reads, in distinct banks; with 33 and 65 bank
  // conflicts took 25% of the step's cycles, but removing them gains only 1%
  // (1403 -> 1384 us at n = 16384, b = 256): the MFMA issue, not LDS, binds.
  static constexpr int PA = sizeof(T) == 8 ? 1 : 2;
  static constexpr int PX = sizeof(T) == 8 ? 1 : 16;
};
template <typename T, int MODE>
__global__ __launch_bounds__(kThreads, 2) void dense_step_kernel(
    int64_t n, const T* __restrict__ A, int64_t lda, const T* __restrict__ X, const T* Z, T* out,
    int ld, int ncols, T p0, T p1, T p2, const T* __restrict__ R, const T* __restrict__ kap,
    int ntn) {
  constexpr int BK = DenseCfg<T>::BK;
  constexpr int BM = DenseCfg<T>::BM, BN = DenseCfg<T>::BN;
  constexpr int AE = BM * BK / kThreads;  // A elements per thread per chunk
  constexpr int XE = BK * BN / kThreads;  // X elements per thread per chunk
  constexpr int WR = BM / 4;              // rows per wave
  constexpr int NA = WR / 16, NB = BN / 16;
  __shared__ T As[2][BM][BK + DenseCfg<T>::PA];
  __shared__ T Xs[2][BK][BN + DenseCfg<T>::PX];
  using Acc = typename MfmaAcc<T>::type;
  const int64_t L = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t ntm = (n + BM - 1) / BM;
  if (L >= ntm * ntn) return;
  const int64_t i0 = (L / ntn) * BM;
  const int c0 = (int)(L % ntn) * BN;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  Acc acc[NA][NB];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = Acc{0, 0, 0, 0};
  // Each thread loads AE consecutive A entries of one row and XE consecutive
  // X entries of one K row, in 16-byte vectors when they are whole and
  // aligned (the chunk offset moves along K).
  using VT = V16<T>;
  constexpr int W = VT::W;
  static_assert(AE % W == 0 && XE % W == 0 && WR % 16 == 0 && BN % 16 == 0, "dense tile shape");
  constexpr int TPA = BK / AE;  // threads per A row
  constexpr int TPX = BN / XE;  // threads per X row
  const int ar = threadIdx.x / TPA, ak = (threadIdx.x % TPA) * AE;
  const int xk = threadIdx.x / TPX, xc = (threadIdx.x % TPX) * XE;
  const int64_t ai = i0 + ar;
  const bool a_vec = ((reinterpret_cast<uintptr_t>(A) & 15) == 0) && (lda % W == 0);
  const bool x_vec = ((reinterpret_cast<uintptr_t>(X) & 15) == 0) && (ld % W == 0) &&
                     (c0 + xc + XE <= ncols);
  // every row of the tile and every column of its X slice exist and both
  // operands are 16-byte aligned: whole chunks are then loaded with no per-lane
  // guard (decided once per wave: loads that lanes take or skip on their own
  // make the wave wait, where the paths join, for the loads still in flight,
  // which would undo the prefetch)
  const bool tile_ok = a_vec && i0 + BM <= n && ((reinterpret_cast<uintptr_t>(X) & 15) == 0) &&
                       (ld % W == 0) && c0 + BN <= ncols;
  T ra2[2][AE], rx2[2][XE];
  auto fetch_to = [&](int64_t k0, T (&ra)[AE], T (&rx)[XE]) {
    const int64_t ka = k0 + ak;
    if (__builtin_amdgcn_readfirstlane((int)(tile_ok && k0 + BK <= n)) != 0) {
      const T* ap = A + ai * lda + ka;
      const T* xp = X + (k0 + xk) * ld + c0 + xc;
#pragma unroll
      for (int q = 0; q < AE / W; ++q) {
        T t[W];
        VT::ld(ap + q * W, t);
#pragma unroll
        for (int w = 0; w < W; ++w) ra[q * W + w] = t[w];
      }
#pragma unroll
      for (int q = 0; q < XE / W; ++q) {
        T t[W];
        VT::ld(xp + q * W, t);
#pragma unroll
        for (int w = 0; w < W; ++w) rx[q * W + w] = t[w];
      }
      return;
    }
    if (a_vec && ai < n && ka + AE <= n) {
      const T* ap = A + ai * lda + ka;
#pragma unroll
      for (int q = 0; q < AE / W; ++q) {
        T t[W];
        VT::ld(ap + q * W, t);
#pragma unroll
        for (int w = 0; w < W; ++w) ra[q * W + w] = t[w];
      }
    } else {
#pragma unroll
      for (int q = 0; q < AE; ++q) ra[q] = (ai < n && ka + q < n) ? A[ai * lda + ka + q] : T(0);
    }
    const int64_t kx = k0 + xk;
    if (x_vec && kx < n) {
      const T* xp = X + kx * ld + c0 + xc;
#pragma unroll
      for (int q = 0; q < XE / W; ++q) {
        T t[W];
        VT::ld(xp + q * W, t);
#pragma unroll
        for (int w = 0; w < W; ++w) rx[q * W + w] = t[w];
      }
    } else {
#pragma unroll
      for (int q = 0; q < XE; ++q) {
        const int c = c0 + xc + q;
        rx[q] = (kx < n && c < ncols) ? X[kx * ld + c] : T(0);
      }
    }
  };
  auto chunk = [&](int buf) {
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kk = ks + (lane >> 4);
      T af[NA], bf[NB];
#pragma unroll
      for (int a = 0; a < NA; ++a) af[a] = As[buf][WR * wave + 16 * a + (lane & 15)][kk];
#pragma unroll
      for (int b = 0; b < NB; ++b) bf[b] = Xs[buf][kk][16 * b + (lane & 15)];
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = mfma16x16x4(af[a], bf[b], acc[a][b]);
    }
  };
  // Chunk c is loaded into register set c % 2 two chunks before its MFMAs
  // and stored to LDS buffer c % 2 at the end of chunk c - 1, so its global
  // loads have two chunks of MFMAs to arrive in (FP64 step at n = 16384,
  // b = 256: 2.77 -> 2.45 ms, 49.6 -> 56.1 TFLOP/s; one chunk ahead left the
  // MFMA pipes idle 37% of the kernel, waves waiting on loads).
  auto put = [&](int buf, const T (&sa)[AE], const T (&sx)[XE]) {
#pragma unroll
    for (int q = 0; q < AE; ++q) As[buf][ar][ak + q] = sa[q];
#pragma unroll
    for (int q = 0; q < XE; ++q) Xs[buf][xk][xc + q] = sx[q];
  };
  fetch_to(0, ra2[0], rx2[0]);
  put(0, ra2[0], rx2[0]);
  if (BK < n) fetch_to(BK, ra2[1], rx2[1]);
  __syncthreads();
  for (int64_t k0 = 0; k0 < n; k0 += 2 * BK) {
    // even chunk: buffer 0, its successor in set 1; fetch chunk + 2 into set 0
    if (k0 + 2 * BK < n) fetch_to(k0 + 2 * BK, ra2[0], rx2[0]);
    chunk(0);
    if (k0 + BK < n) put(1, ra2[1], rx2[1]);
    __syncthreads();
    if (k0 + BK >= n) break;
    // odd chunk: buffer 1, its successor in set 0; fetch chunk + 2 into set 1
    if (k0 + 3 * BK < n) fetch_to(k0 + 3 * BK, ra2[1], rx2[1]);
    chunk(1);
    if (k0 + 2 * BK < n) put(0, ra2[0], rx2[0]);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t i = i0 + WR * wave + 16 * a + MfmaAcc<T>::row(j, lane);
        const int c = c0 + 16 * b + (lane & 15);
        if (i >= n || c >= ncols) continue;
        const int64_t off = i * ld + c;
        const T v = acc[a][b][j];
        T o;
        if constexpr (MODE == kApply) o = p0 * v;
        else if constexpr (MODE == kApplyBeta) o = fma(p0, v, p1 * out[off]);
        else if constexpr (MODE == kChebFirst) o = p2 * (p1 * fma(p0, v, X[off]));
        else if constexpr (MODE == kCheb) o = p2 * fma(p1, fma(p0, v, X[off]), -Z[off]);
        else o = p2 * fma(kap[c], R[off], fma(p1, fma(p0, v, X[off]), -Z[off]));
        out[off] = o;
      }
}

// One column of a dense operator (the dense Lanczos matrix-vector products):
// out = p0 A x (+ p1 out with kApplyBeta).  A wave owns a row: its lanes read
// 16-byte pieces of the row in order and sum them, then a fixed shuffle tree
// adds the lanes.
template <int MODE>
__global__ __launch_bounds__(kThreads) void dense_gemv_kernel(int64_t n,
                                                              const double* __restrict__ A,
                                                              int64_t lda,
                                                              const double* __restrict__ x,
                                                              double* out, double p0, double p1) {
  const int lane = threadIdx.x & 63;
  const int64_t i = blockIdx.x * (int64_t)(kThreads / 64) + (threadIdx.x >> 6);
  if (i >= n) return;  // uniform across the wave
  const double* a = A + i * lda;
  const bool vec = ((reinterpret_cast<uintptr_t>(a) & 15) == 0) &&
                   ((reinterpret_cast<uintptr_t>(x) & 15) == 0);
  double s0 = 0.0, s1 = 0.0;
  int64_t k = 2 * (int64_t)lane;
  if (vec) {
    for (; k + 1 < n; k += 128) {
      const dvec2 av = *reinterpret_cast<const dvec2*>(a + k);
      const dvec2 xv = *reinterpret_cast<const dvec2*>(x + k);
      s0 = fma(av.x, xv.x, s0);
      s1 = fma(av.y, xv.y, s1);
    }
    if (k < n) s0 = fma(a[k], x[k], s0);
  } else {
    for (; k < n; k += 128) {
      s0 = fma(a[k], x[k], s0);
      if (k + 1 < n) s1 = fma(a[k + 1], x[k + 1], s1);
    }
  }
  double s = s0 + s1;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) {
    if constexpr (MODE == kApplyBeta) out[i] = fma(p0, s, p1 * out[i]);
    else out[i] = p0 * s;
  }
}

// ------------------------------------------------------ column reductions
// Threads are laid out as (row slot, column): CW columns per pass.
constexpr int kColRows = 512;  // rows per partial block
// Rows whose loads a column-reduction thread issues together before adding
// them in row order (the sums are those of the one-row loop): at n ~ 5e4,
// b = 24 a thread walks ~50 rows, one load latency each when taken singly.
constexpr int kColDepth = 8;

template <typename T, bool RESID>
__global__ __launch_bounds__(kThreads) void col_sumsq_kernel(int64_t n, const T* __restrict__ V,
                                                             const T* __restrict__ Wm, int ld,
                                                             int nc, const double* theta,
                                                             double* __restrict__ P) {
  __shared__ double red[kThreads];
  const int64_t blk = blockIdx.x;
  const int cw = min(nc, kThreads);  // columns per pass
  const int slots = kThreads / cw;
  const int c_in = threadIdx.x % cw;
  const int slot = threadIdx.x / cw;
  const int64_t i0 = blk * kColRows;
  const int64_t i1 = min(n, i0 + kColRows);
  for (int cbase = 0; cbase < nc; cbase += cw) {
    const int c = cbase + c_in;
    double s = 0.0;
    if (slot < slots && c < nc) {
      const double th = RESID ? theta[c] : 0.0;
      int64_t i = i0 + slot;
      // kColDepth rows' loads in flight, summed in row order as one at a time
      for (; i + (kColDepth - 1) * slots < i1; i += kColDepth * slots) {
        double v[kColDepth], w[kColDepth];
#pragma unroll
        for (int k = 0; k < kColDepth; ++k) {
          v[k] = double(V[(i + k * slots) * ld + c]);
          if (RESID) w[k] = double(Wm[(i + k * slots) * ld + c]);
        }
#pragma unroll
        for (int k = 0; k < kColDepth; ++k) {
          double x = v[k];
          if (RESID) x = th * x - w[k];
          s += x * x;
        }
      }
      for (; i < i1; i += slots) {
        double v = double(V[i * ld + c]);
        if (RESID) v = th * v - double(Wm[i * ld + c]);
        s += v * v;
      }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    if (slot == 0 && c < nc) {
      double t = red[threadIdx.x];
      for (int q = 1; q < slots; ++q) t += red[threadIdx.x + q * cw];
      P[blk * nc + c] = t;
    }
    __syncthreads();
  }
}

// Residual and norm sums together: P[blk*2nc + c] = sum_i |theta_c V - W|^2,
// P[blk*2nc + nc + c] = sum_i |V|^2 over the block's rows.
// TH: const double* (theta in device memory) or ThetaArgs (in the arguments)
template <typename T, typename TH>
__global__ __launch_bounds__(kThreads) void resid_norm_kernel(int64_t n, const T* __restrict__ V,
                                                              const T* __restrict__ Wm, int ld,
                                                              int nc, const TH theta,
                                                              double* __restrict__ P) {
  __shared__ double red[2 * kThreads];
  const int64_t blk = blockIdx.x;
  const int cw = min(nc, kThreads);
  const int slots = kThreads / cw;
  const int c_in = threadIdx.x % cw;
  const int slot = threadIdx.x / cw;
  const int64_t i0 = blk * kColRows;
  const int64_t i1 = min(n, i0 + kColRows);
  for (int cbase = 0; cbase < nc; cbase += cw) {
    const int c = cbase + c_in;
    double s = 0.0, q = 0.0;
    if (slot < slots && c < nc) {
      const double th = theta[c];
      int64_t i = i0 + slot;
      for (; i + (kColDepth - 1) * slots < i1; i += kColDepth * slots) {
        double v[kColDepth], w[kColDepth];
#pragma unroll
        for (int k = 0; k < kColDepth; ++k) {
          v[k] = double(V[(i + k * slots) * ld + c]);
          w[k] = double(Wm[(i + k * slots) * ld + c]);
        }
#pragma unroll
        for (int k = 0; k < kColDepth; ++k) {
          const double r = th * v[k] - w[k];
          s += r * r;
          q += v[k] * v[k];
        }
      }
      for (; i < i1; i += slots) {
        const double v = double(V[i * ld + c]);
        const double r = th * v - double(Wm[i * ld + c]);
        s += r * r;
        q += v * v;
      }
    }
    red[threadIdx.x] = s;
    red[kThreads + threadIdx.x] = q;
    __syncthreads();
    if (slot == 0 && c < nc) {
      double t = red[threadIdx.x], u = red[kThreads + threadIdx.x];
      for (int k = 1; k < slots; ++k) {
        t += red[threadIdx.x + k * cw];
        u += red[kThreads + threadIdx.x + k * cw];
      }
      P[blk * 2 * nc + c] = t;
      P[blk * 2 * nc + nc + c] = u;
    }
    __syncthreads();
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void normalize_cols_kernel(int64_t n, T* X, int ld, int nc,
                                                                  const double* __restrict__ sums) {
  const int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  const int64_t i = e / nc;
  const int c = (int)(e % nc);
  if (i >= n) return;
  const double s = 1.0 / sqrt(fabs(sums[c]));
  X[i * ld + c] = T(double(X[i * ld + c]) * s);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void copy_cols_kernel(int64_t n, const T* __restrict__ src,
                                                             int lds, int s0, T* __restrict__ dst,
                                                             int ldd, int d0, int cnt) {
  const int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  const int64_t i = e / cnt;
  const int j = (int)(e % cnt);
  if (i >= n) return;
  dst[i * ldd + d0 + j] = src[i * lds + s0 + j];
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double sym_draw(uint64_t key, uint64_t idx) {
  const uint64_t z = mix64(key + (idx + 1) * 0x9E3779B97F4A7C15ULL);
  return -1.0 + 2.0 * (double(z >> 11) * 0x1.0p-53);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void randomize_cols_kernel(int64_t n, T* X, int ld, int c0,
                                                                  int cnt, uint64_t key,
                                                                  uint64_t counter) {
  // thread per (row, column): column-fastest so stores are coalesced.
  const int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  const int64_t i = e / cnt;
  const int j = (int)(e % cnt);
  if (i >= n) return;
  X[i * ld + c0 + j] = T(sym_draw(key, counter + (uint64_t)j * (uint64_t)n + (uint64_t)i));
}

template <typename T>
__global__ __launch_bounds__(kThreads) void fill_cols_kernel(int64_t n, T* X, int ld, int c0,
                                                             int cnt, T value) {
  const int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  const int64_t i = e / cnt;
  const int j = (int)(e % cnt);
  if (i >= n) return;
  X[i * ld + c0 + j] = value;
}

// ------------------------------------------------ mixed-precision filter
// Start of the correction recurrence X_j = u_j(theta_c) v + E_j: from
// V (FP64) and AV (FP64), per column c < nc with cp[c] = theta_c and
// cp[nc + c] = 1 (Ritz column) or 0 (plain column):
//   Ritz : R = AV - theta V, E0 = 0, E1 = e1 * R
//   plain: R = 0,            E0 = V, E1 = g2*(g1*AV + V)
// E0, E1, R are FP32 blocks of the same ld.
__global__ __launch_bounds__(kThreads) void mixed_start_kernel(
    int64_t n, const double* __restrict__ V, const double* __restrict__ AV, int ld, int nc,
    const double* __restrict__ cp, double e1, double g1, double g2, float* __restrict__ E0,
    float* __restrict__ E1, float* __restrict__ R) {
  const int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  const int64_t i = e / nc;
  const int c = (int)(e % nc);
  if (i >= n) return;
  const int64_t k = i * ld + c;
  float e0, e1v, r;
  mixed_start_elem(V[k], AV[k], cp[c], cp[nc + c], e1, g1, g2, e0, e1v, r);
  R[k] = r;
  E0[k] = e0;
  E1[k] = e1v;
}

// V <- s * (um[c] * V + E), c < nc (E FP32).
__global__ __launch_bounds__(kThreads) void mixed_combine_kernel(int64_t n, double* V, int ld,
                                                                 int nc, const double* __restrict__ um,
                                                                 double s, const float* __restrict__ E) {
  const int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  const int64_t i = e / nc;
  const int c = (int)(e % nc);
  if (i >= n) return;
  const int64_t k = i * ld + c;
  V[k] = s * fma(um[c], V[k], double(E[k]));
}

__global__ __launch_bounds__(kThreads) void to_f32_kernel(int64_t n, const double* __restrict__ x,
                                                          float* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  if (i < n) y[i] = float(x[i]);
}

// ------------------------------------------------------- complex helpers
// Complex block columns are pairs of real columns (re, im).  Start vectors:
// complex column c0+j, row i takes draws counter + 2(j*n + i) + {0 re, 1 im}.
__global__ __launch_bounds__(kThreads) void randomize_cplx_kernel(int64_t n, double* X, int ld,
                                                                  int c0, int cnt, uint64_t key,
                                                                  uint64_t counter) {
  const int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  const int64_t i = e / (2 * cnt);
  const int r = (int)(e % (2 * cnt));
  if (i >= n) return;
  const int j = r >> 1;
  const int part = r & 1;
  X[i * ld + 2 * (c0 + j) + part] =
      sym_draw(key, counter + 2 * ((uint64_t)j * (uint64_t)n + (uint64_t)i) + (uint64_t)part);
}

// s[2c] = s[2c+1] = s[2c] + s[2c+1], c < nc: squared norms of complex columns
// from those of their real and imaginary parts.
__global__ void pair_sums_kernel(double* s, int nc) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const double t = s[2 * c] + s[2 * c + 1];
  s[2 * c] = t;
  s[2 * c + 1] = t;
}

// In place, per 2x2 block: the real Gram G~ = Y~^T V~ of interleaved complex
// blocks (2f x 2b, column-major, ld 2f) becomes the real embedding of
// C = Y^H V, so that V~ C~ is the interleaved form of V C:
//   C = (G~[2a,2c] + G~[2a+1,2c+1]) + i (G~[2a,2c+1] - G~[2a+1,2c]),
//   C~[2a,2c] = C~[2a+1,2c+1] = Re C, C~[2a,2c+1] = Im C, C~[2a+1,2c] = -Im C.
__global__ void herm_embed_kernel(double* G, int f, int b) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= f * b) return;
  const int a = e % f;
  const int c = e / f;
  const int64_t ld = 2 * f;
  double* g00 = G + 2 * a + (int64_t)(2 * c) * ld;
  double* g10 = g00 + 1;
  double* g01 = G + 2 * a + (int64_t)(2 * c + 1) * ld;
  double* g11 = g01 + 1;
  const double re = *g00 + *g11;
  const double im = *g01 - *g10;
  *g00 = re;
  *g11 = re;
  *g01 = im;
  *g10 = -im;
}

// ----------------------------------------------------------------- Lanczos
// rows per partial block of the Lanczos kernels: 512 below 2^20 rows (more
// blocks for small operators), 2048 from there (fewer partials to reduce)
inline int64_t lz_rows(int64_t n) { return n >= (int64_t(1) << 20) ? 2048 : 512; }

__device__ __forceinline__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// v = v/||v||, w = v, v = 0 (LanczosMaxMinFinder.h:272-279) -- two phases.
__global__ __launch_bounds__(kThreads) void lz_sumsq_kernel(int64_t n, const double* v, double* P, int64_t rpb) {
  __shared__ double red[kThreads];
  const int64_t i0 = blockIdx.x * rpb;
  const int64_t i1 = min(n, i0 + rpb);
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) s += v[i] * v[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) P[blockIdx.x] = s;
}
__global__ __launch_bounds__(kThreads) void lz_finalize_kernel(const double* P, int64_t nblk,
                                                               double* out, int take_sqrt) {
  __shared__ double red[kThreads];
  // fixed-order: each thread sums a strided subset, then a tree.
  double s = 0.0;
  for (int64_t b = threadIdx.x; b < nblk; b += kThreads) s += P[b];
  s = block_sum(s, red);
  if (threadIdx.x == 0) *out = take_sqrt ? sqrt(fabs(s)) : s;
}
__global__ __launch_bounds__(kThreads) void lz_init_kernel(int64_t n, double* v, double* w,
                                                           const double* nrm) {
  const int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  if (i >= n) return;
  const double s = 1.0 / *nrm;
  w[i] = v[i] * s;
  v[i] = 0.0;
}
// if |beta| > 1e-14: w <- v/beta, v <- -beta*w_old (LanczosMaxMinFinder.h:329-339)
__global__ __launch_bounds__(kThreads) void lz_swap_kernel(int64_t n, double* v, double* w,
                                                           const double* beta_prev) {
  const double b = *beta_prev;
  if (!(fabs(b) > 1.0e-14)) return;
  const int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  if (i >= n) return;
  const double t = w[i];
  w[i] = v[i] * (1.0 / b);
  v[i] = t * (-b);
}
// v += A w; partial dot(w, v)
__global__ __launch_bounds__(kThreads) void lz_apply_dot_kernel(int64_t n,
                                                                const int* __restrict__ rowptr,
                                                                const int* __restrict__ col,
                                                                const double* __restrict__ val,
                                                                const double* __restrict__ w,
                                                                double* v, double* P, int64_t rpb) {
  __shared__ double red[kThreads];
  const int64_t i0 = blockIdx.x * rpb;
  const int64_t i1 = min(n, i0 + rpb);
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
    double a = 0.0;
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) a = fma(val[p], w[col[p]], a);
    const double vn = v[i] + a;
    v[i] = vn;
    s = fma(w[i], vn, s);
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) P[blockIdx.x] = s;
}
// complex: v += A w (interleaved re, im), partials of Re and Im of sum conj(w) v
__global__ __launch_bounds__(kThreads) void lz_apply_dot_c_kernel(
    int64_t n, const int* __restrict__ rowptr, const int* __restrict__ col,
    const double* __restrict__ val, const double* __restrict__ w, double* v, double* P,
    int64_t nblk, int64_t rpb) {
  __shared__ double red[kThreads];
  const int64_t i0 = blockIdx.x * rpb;
  const int64_t i1 = min(n, i0 + rpb);
  double sr = 0.0, si = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
    double ar = 0.0, ai = 0.0;
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const double vr = val[2 * (int64_t)p], vi = val[2 * (int64_t)p + 1];
      const double xr = w[2 * (int64_t)col[p]], xi = w[2 * (int64_t)col[p] + 1];
      ar = fma(vr, xr, fma(-vi, xi, ar));
      ai = fma(vr, xi, fma(vi, xr, ai));
    }
    const double nr = v[2 * i] + ar, ni = v[2 * i + 1] + ai;
    v[2 * i] = nr;
    v[2 * i + 1] = ni;
    const double wr = w[2 * i], wi = w[2 * i + 1];
    sr = fma(wr, nr, fma(wi, ni, sr));
    si = fma(wr, ni, fma(-wi, nr, si));
  }
  sr = block_sum(sr, red);
  si = block_sum(si, red);
  if (threadIdx.x == 0) {
    P[blockIdx.x] = sr;
    P[nblk + blockIdx.x] = si;
  }
}

// partial sum w . v (the dense operator's Lanczos alpha)
__global__ __launch_bounds__(kThreads) void lz_dot_kernel(int64_t n, const double* w,
                                                          const double* v, double* P, int64_t rpb) {
  __shared__ double red[kThreads];
  const int64_t i0 = blockIdx.x * rpb;
  const int64_t i1 = min(n, i0 + rpb);
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) s = fma(w[i], v[i], s);
  s = block_sum(s, red);
  if (threadIdx.x == 0) P[blockIdx.x] = s;
}

// v += (-alpha) w; partial |v|^2
__global__ __launch_bounds__(kThreads) void lz_axpy_norm_kernel(int64_t n, const double* w,
                                                                double* v, const double* alpha,
                                                                double* P, int64_t rpb) {
  __shared__ double red[kThreads];
  const double a = -*alpha;
  const int64_t i0 = blockIdx.x * rpb;
  const int64_t i1 = min(n, i0 + rpb);
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
    const double vn = fma(w[i], a, v[i]);
    v[i] = vn;
    s = fma(vn, vn, s);
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) P[blockIdx.x] = s;
}

// The block's partial sum goes to P[blockIdx.x]; the last block to finish
// (counted on *cnt, which it resets) adds the partials in block order -- the
// order lz_finalize_kernel uses -- and writes the total (or its square root).
// No block waits for another.
__device__ void last_block_total(double s, double* P, unsigned* cnt, double* out, int take_sqrt,
                                 double* red) {
  __shared__ int am_last;
  if (threadIdx.x == 0) {
    P[blockIdx.x] = s;
    __threadfence();
    const unsigned t = atomicAdd(cnt, 1u);
    am_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return;
  __threadfence();
  double a = 0.0;
  for (int64_t b = threadIdx.x; b < (int64_t)gridDim.x; b += kThreads) a += P[b];
  a = block_sum(a, red);
  if (threadIdx.x == 0) {
    *out = take_sqrt ? sqrt(fabs(a)) : a;
    *cnt = 0u;
  }
}

// One Lanczos step on a ping-pong pair of vectors (LanczosMaxMinFinder.h:
// 300-345), for real CSR operators.  The previous step's swap (when
// |beta_{k-1}| > 1e-14: w' = v / beta, v' = -beta w) is folded into the
// operator application, whose gathers read v / beta directly; the result
// goes to the other pair (vn = v' + A w', wn = w'), and alpha_k = w' . vn.
// Same arithmetic, in the same order, as lz_swap + lz_apply_dot + finalize.
__global__ __launch_bounds__(kThreads) void lz_step_kernel(
    int64_t n, const int* __restrict__ rowptr, const int* __restrict__ col,
    const double* __restrict__ val, const double* __restrict__ v, const double* __restrict__ w,
    double* __restrict__ vn, double* __restrict__ wn, const double* beta_prev, double* P,
    unsigned* cnt, double* alpha_k, int64_t rpb) {
  __shared__ double red[kThreads];
  const double b = beta_prev ? *beta_prev : 0.0;
  const bool sw = beta_prev != nullptr && fabs(b) > 1.0e-14;
  const double rb = sw ? 1.0 / b : 1.0;
  const int64_t i0 = blockIdx.x * rpb;
  const int64_t i1 = min(n, i0 + rpb);
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
    double a = 0.0;
    if (sw) {
      for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) a = fma(val[p], v[col[p]] * rb, a);
    } else {
      for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) a = fma(val[p], w[col[p]], a);
    }
    const double wi = sw ? v[i] * rb : w[i];
    const double vi = sw ? w[i] * (-b) : v[i];
    const double vnew = vi + a;
    vn[i] = vnew;
    wn[i] = wi;
    s = fma(wi, vnew, s);
  }
  s = block_sum(s, red);
  last_block_total(s, P, cnt, alpha_k, 0, red);
}

// v -= alpha w; beta_k = ||v|| (lz_axpy_norm + finalize in one launch)
__global__ __launch_bounds__(kThreads) void lz_axpy_beta_kernel(int64_t n, const double* w,
                                                                double* v, const double* alpha,
                                                                double* P, unsigned* cnt,
                                                                double* beta_k, int64_t rpb) {
  __shared__ double red[kThreads];
  const double a = -*alpha;
  const int64_t i0 = blockIdx.x * rpb;
  const int64_t i1 = min(n, i0 + rpb);
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
    const double vn = fma(w[i], a, v[i]);
    v[i] = vn;
    s = fma(vn, vn, s);
  }
  s = block_sum(s, red);
  last_block_total(s, P, cnt, beta_k, 1, red);
}

inline unsigned grid_for(int64_t work) {
  return (unsigned)((work + kThreads - 1) / kThreads);
}

}  // namespace

// ============================================================== launchers
template <typename T>
void launch_csr_apply(int64_t n, const int* rowptr, const int* col, const T* val, const T* X,
                      T* out, int ld, int ncols, T alpha, T beta, hipStream_t st,
                      const CsrSched* sc, bool cx) {
  if (n <= 0 || ncols <= 0) return;
  if (beta == T(0))
    csr_launch<T, kApply>(n, rowptr, col, val, X, nullptr, out, ld, ncols, alpha, T(0), T(0), st, sc,
                          nullptr, nullptr, cx);
  else
    csr_launch<T, kApplyBeta>(n, rowptr, col, val, X, nullptr, out, ld, ncols, alpha, beta, T(0), st,
                              sc, nullptr, nullptr, cx);
}

void launch_csr_apply_start(int64_t n, const int* rowptr, const int* col, const double* val,
                            const double* V, int ld, int ncols, const MixedArgs& ms,
                            hipStream_t st, const CsrSched* sc, bool cx) {
  if (n <= 0 || ncols <= 0) return;
  if (ms.nc != ncols || !ms.cp || !ms.e0 || !ms.e1 || !ms.r)
    throw Error(1, "fused start: coefficient table and FP32 blocks must cover the call's columns");
  csr_launch<double, kApplyStart>(n, rowptr, col, val, V, nullptr, nullptr, ld, ncols, 1.0, 0.0,
                                  0.0, st, sc, nullptr, nullptr, cx, ms);
}

template <typename T>
void launch_cheb_step(int64_t n, const int* rowptr, const int* col, const T* val, const T* X1,
                      const T* X0, T* out, int ld, int ncols, T g1, T g2, T s, bool first,
                      hipStream_t st, int panel, const CsrSched* sc, bool cx) {
  if (n <= 0 || ncols <= 0) return;
  // Column panels: each launch streams the CSR arrays once and touches only
  // `panel` columns of every row, which shrinks the per-plane footprint of
  // the gathered block so that neighbour rows stay in L2 between uses.
  if (panel <= 0 || panel >= ncols) panel = ncols;
  if (panel % 4 != 0) panel = round_up(panel, 4);
  for (int c0 = 0; c0 < ncols; c0 += panel) {
    const int pc = std::min(panel, ncols - c0);
    if (first)
      csr_launch<T, kChebFirst>(n, rowptr, col, val, X1 + c0, nullptr, out + c0, ld, pc, g1, g2, s,
                                st, sc, nullptr, nullptr, cx);
    else
      csr_launch<T, kCheb>(n, rowptr, col, val, X1 + c0, X0 + c0, out + c0, ld, pc, g1, g2, s, st,
                           sc, nullptr, nullptr, cx);
  }
}

void launch_cheb_step_src(int64_t n, const int* rowptr, const int* col, const float* val,
                          const float* X1, const float* X0, float* out, const float* R,
                          const float* kap, int ld, int ncols, float g1, float g2, float s,
                          hipStream_t st, int panel, const CsrSched* sc, bool cx) {
  if (n <= 0 || ncols <= 0) return;
  if (panel <= 0 || panel >= ncols) panel = ncols;
  if (panel % 4 != 0) panel = round_up(panel, 4);
  for (int c0 = 0; c0 < ncols; c0 += panel) {
    const int pc = std::min(panel, ncols - c0);
    csr_launch<float, kChebSrc>(n, rowptr, col, val, X1 + c0, X0 + c0, out + c0, ld, pc, g1, g2, s,
                                st, sc, R + c0, kap + c0, cx);
  }
}

void launch_cheb_step_src_end(int64_t n, const int* rowptr, const int* col, const float* val,
                              const float* X1, const float* X0, const float* R, const float* kap,
                              int ld, int ncols, float g1, float g2, const MixedArgs& end,
                              hipStream_t st, int panel, const CsrSched* sc, bool cx) {
  if (n <= 0 || ncols <= 0) return;
  if (!end.v_end || !end.um) throw Error(1, "fused recombination: V and um are required");
  if (panel <= 0 || panel >= ncols) panel = ncols;
  if (panel % 4 != 0) panel = round_up(panel, 4);
  for (int c0 = 0; c0 < ncols; c0 += panel) {
    const int pc = std::min(panel, ncols - c0);
    MixedArgs e = end;
    e.v_end = end.v_end + c0;
    e.um = end.um + c0;
    csr_launch<float, kChebSrcEnd>(n, rowptr, col, val, X1 + c0, X0 + c0, nullptr, ld, pc, g1, g2,
                                   1.0f, st, sc, R + c0, kap + c0, cx, e);
  }
}

void launch_mixed_start(int64_t n, const double* V, const double* AV, int ld, int nc,
                        const double* cp, double e1, double g1, double g2, float* E0, float* E1,
                        float* R, hipStream_t st) {
  if (n <= 0 || nc <= 0) return;
  const int64_t tot = n * nc;
  hipLaunchKernelGGL(mixed_start_kernel, dim3((unsigned)((tot + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, st, n, V, AV, ld, nc, cp, e1, g1, g2, E0, E1, R);
  RCG_HIP(hipGetLastError());
}

void launch_mixed_combine(int64_t n, double* V, int ld, int nc, const double* um, double s,
                          const float* E, hipStream_t st) {
  if (n <= 0 || nc <= 0) return;
  const int64_t tot = n * nc;
  hipLaunchKernelGGL(mixed_combine_kernel, dim3((unsigned)((tot + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, st, n, V, ld, nc, um, s, E);
  RCG_HIP(hipGetLastError());
}

void launch_to_f32(int64_t n, const double* x, float* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(to_f32_kernel, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads),
                     0, st, n, x, y);
  RCG_HIP(hipGetLastError());
}

template <typename T>
void launch_dense_step(int64_t n, const T* A, int64_t lda, const T* X, const T* Z, T* out, int ld,
                       int ncols, T p0, T p1, T p2, int mode, const T* R, const T* kap,
                       hipStream_t st) {
  if (n <= 0 || ncols <= 0) return;
  if constexpr (std::is_same<T, double>::value) {
    if (ncols == 1 && ld == 1 && (mode == kApply || mode == kApplyBeta)) {
      const dim3 g1((unsigned)((n + kThreads / 64 - 1) / (kThreads / 64)));
      if (mode == kApply)
        hipLaunchKernelGGL(dense_gemv_kernel<kApply>, g1, dim3(kThreads), 0, st, n, A, lda, X,
                           out, p0, p1);
      else
        hipLaunchKernelGGL(dense_gemv_kernel<kApplyBeta>, g1, dim3(kThreads), 0, st, n, A, lda, X,
                           out, p0, p1);
      RCG_HIP(hipGetLastError());
      return;
    }
  }
  const int ntn = (ncols + DenseCfg<T>::BN - 1) / DenseCfg<T>::BN;
  const int64_t tiles = ((n + DenseCfg<T>::BM - 1) / DenseCfg<T>::BM) * ntn;
  const dim3 grid((unsigned)round_up(tiles, kXcds));
#define RCG_DENSE(M)                                                                           \
  hipLaunchKernelGGL((dense_step_kernel<T, M>), grid, dim3(kThreads), 0, st, n, A, lda, X, Z, out, \
                     ld, ncols, p0, p1, p2, R, kap, ntn)
  switch (mode) {
    case kApply: RCG_DENSE(kApply); break;
    case kApplyBeta: RCG_DENSE(kApplyBeta); break;
    case kChebFirst: RCG_DENSE(kChebFirst); break;
    case kCheb: RCG_DENSE(kCheb); break;
    case kChebSrc: RCG_DENSE(kChebSrc); break;
    default: throw Error(1, "unknown dense step mode");
  }
#undef RCG_DENSE
  RCG_HIP(hipGetLastError());
}

template <typename T>
size_t gram_ws_elems(int64_t n, int bx, int by) {
  int64_t rpb;
  const int64_t nblk = gram_blocks<T>(n, bx, by, &rpb);
  return (size_t)nblk * bx * by;
}

// Upper bound of gram_ws_elems over every bx, by <= bmax: a launch has at
// most 1024/ntiles + 1 row blocks (and at most ceil(n/kRb)), each holding at
// most ntiles 64x64 tiles.
size_t gram_ws_bound(int64_t n, int bmax) {
  const int64_t t = (bmax + kTile - 1) / kTile;
  const int64_t by_tiles = (1024 + t * t) * (int64_t)kTile * kTile;
  const int64_t by_rows = ((n + kRb - 1) / kRb) * (int64_t)bmax * bmax;
  return (size_t)std::max<int64_t>(1, std::min(by_tiles, by_rows));
}

// 16x16 subtiles per side of a one-tile Gram launch
inline int gram_nt(int bx, int by) { return (std::max(bx, by) + 15) / 16; }
template <int MODE>
void gram_one_tile(int nt, dim3 grid, hipStream_t st, int64_t n, const double* X, int ldx, int bx,
                   const double* Y, int ldy, int by, double* P, int64_t rpb, int nta, int ntiles,
                   int64_t nblk) {
#define RCG_GRAM_NT(NT)                                                                          \
  case NT:                                                                                       \
    hipLaunchKernelGGL((gram_mfma_kernel<MODE, NT>), grid, dim3(kThreads), 0, st, n, X, ldx, bx, \
                       Y, ldy, by, P, rpb, nta, ntiles, nblk);                                   \
    break;
  switch (nt) {
    RCG_GRAM_NT(1)
    RCG_GRAM_NT(2)
    RCG_GRAM_NT(3)
    RCG_GRAM_NT(4)
    default: throw Error(1, "one-tile Gram launches take at most 64 columns");
  }
#undef RCG_GRAM_NT
}

template <typename T>
void launch_gram(int64_t n, const T* X, int ldx, int bx, const T* Y, int ldy, int by, T* G,
                 T* ws, size_t ws_elems, hipStream_t st, bool lower_only) {
  if (bx <= 0 || by <= 0) return;
  if (n <= 0) {
    RCG_HIP(hipMemsetAsync(G, 0, sizeof(T) * bx * by, st));
    return;
  }
  int64_t rpb;
  int64_t nblk = gram_blocks<T>(n, bx, by, &rpb);
  if ((size_t)nblk * bx * by > ws_elems) throw Error(3, "gram workspace too small");
  const int nta = (bx + kTile - 1) / kTile;
  const int ntb = (by + kTile - 1) / kTile;
  const int ntiles = nta * ntb;
  if constexpr (std::is_same<T, double>::value) {
    const bool sym = X == Y && bx == by && ldx == ldy;
    if (ntiles > 1) {
      // LDS-staged tiles; the launched tiles times the row blocks fill two
      // workgroups per CU of a 256-CU part once
      const bool low = !sym && lower_only && bx == by;
      const int nl = (sym || low) ? nta * (nta + 1) / 2 : ntiles;
      const int64_t target = std::max<int64_t>(1, 512 / nl);
      rpb = round_up((n + target - 1) / target, kRb);
      rpb = std::max<int64_t>(rpb, round_up(4 * (int64_t)std::max(bx, by), kRb));
      nblk = (n + rpb - 1) / rpb;
      if ((size_t)nblk * bx * by > ws_elems) throw Error(3, "gram workspace too small");
      const dim3 grid((unsigned)round_up(nblk * nl, kXcds));
      if (sym)
        hipLaunchKernelGGL(gram_lds_kernel<kGramSym>, grid, dim3(kThreads), 0, st, n, X, ldx, bx,
                           Y, ldy, by, ws, rpb, nta, nl, nblk);
      else if (low)
        hipLaunchKernelGGL(gram_lds_kernel<kGramLower>, grid, dim3(kThreads), 0, st, n, X, ldx, bx,
                           Y, ldy, by, ws, rpb, nta, nl, nblk);
      else
        hipLaunchKernelGGL(gram_lds_kernel<kGramFull>, grid, dim3(kThreads), 0, st, n, X, ldx, bx,
                           Y, ldy, by, ws, rpb, nta, nl, nblk);
    } else if (sym) {
      gram_one_tile<kGramDiag>(gram_nt(bx, by), dim3((unsigned)round_up(nblk, kXcds)), st, n, X,
                               ldx, bx, Y, ldy, by, ws, rpb, nta, 1, nblk);
    } else {
      gram_one_tile<kGramFull>(gram_nt(bx, by), dim3((unsigned)round_up(nblk * ntiles, kXcds)), st,
                               n, X, ldx, bx, Y, ldy, by, ws, rpb, nta, ntiles, nblk);
    }
  } else {
    hipLaunchKernelGGL(gram_partial_kernel<T>, dim3((unsigned)nblk, ntiles), dim3(kThreads), 0, st,
                       n, X, ldx, bx, Y, ldy, by, ws, rpb, ntiles);
  }
  RCG_HIP(hipGetLastError());
  const int64_t m = (int64_t)bx * by;
  hipLaunchKernelGGL((reduce_partials_kernel<T, T>), dim3(reduce_grid(m)), dim3(kThreads), 0, st, ws,
                     nblk, m, G);
  RCG_HIP(hipGetLastError());
}

void launch_gram_pair(int64_t n, const double* X, int ldx, const double* Y, int ldy, int b,
                      double* G, double* ws, size_t ws_elems, hipStream_t st) {
  if (b <= 0) return;
  if (b > kTile) throw Error(1, "paired Gram products take at most 64 columns");
  if (n <= 0) {
    RCG_HIP(hipMemsetAsync(G, 0, sizeof(double) * 2 * b * b, st));
    return;
  }
  // the row blocks of the single launches, so that the partials and their
  // sums are the same
  int64_t rpb;
  const int64_t nblk = gram_blocks<double>(n, b, b, &rpb);
  if ((size_t)(2 * nblk) * b * b > ws_elems) throw Error(3, "gram workspace too small");
  gram_one_tile<kGramPair>(gram_nt(b, b), dim3((unsigned)round_up(2 * nblk, kXcds)), st, n, X, ldx,
                          b, Y, ldy, b, ws, rpb, 1, 2, nblk);
  RCG_HIP(hipGetLastError());
  const int64_t m = 2 * (int64_t)b * b;
  hipLaunchKernelGGL((reduce_partials_kernel<double, double>), dim3(reduce_grid(m)), dim3(kThreads),
                     0, st, ws, nblk, m, G);
  RCG_HIP(hipGetLastError());
}

template <typename T>
void launch_tall_gemm(int64_t n, const T* X, int ldx, int bx, const T* C, int ldc, int by,
                      T alpha, T beta, const T* Z, T* out, int ldo, hipStream_t st) {
  if (n <= 0 || by <= 0) return;
  const int64_t nrb = (n + kTile - 1) / kTile;
  const int64_t grid = round_up(nrb, kXcds);
  const unsigned ct = (unsigned)((by + kTile - 1) / kTile);
  if constexpr (std::is_same<T, double>::value) {
    if (bx > kTile && bx <= 4 * kTile) {
      // C slice resident in LDS (66.5, 100 or 133 KB) and about 300 registers
      // per lane: one workgroup per CU, persistent over 128-row tiles
      const int kch = (bx + kTile - 1) / kTile;
      const int64_t nrt = (n + 2 * kTile - 1) / (2 * kTile);
      const int64_t nstripe = std::max<int64_t>(1, std::min<int64_t>(nrt, 256 / (int64_t)ct));
      const dim3 grid((unsigned)round_up(nstripe * ct, kXcds));
      if (kch == 2)
        hipLaunchKernelGGL(tall_gemm_wide_kernel<2>, grid, dim3(kThreads), 0, st, n, X, ldx, bx, C,
                           ldc, by, alpha, beta, Z, out, ldo, nrt, (int)ct, nstripe);
      else if (kch == 3)
        hipLaunchKernelGGL(tall_gemm_wide_kernel<3>, grid, dim3(kThreads), 0, st, n, X, ldx, bx, C,
                           ldc, by, alpha, beta, Z, out, ldo, nrt, (int)ct, nstripe);
      else
        hipLaunchKernelGGL(tall_gemm_wide_kernel<4>, grid, dim3(kThreads), 0, st, n, X, ldx, bx, C,
                           ldc, by, alpha, beta, Z, out, ldo, nrt, (int)ct, nstripe);
      RCG_HIP(hipGetLastError());
      return;
    }
    // persistent over row tiles: as many workgroups as fit at once on a
    // 256-CU part (four per CU, three with the registers that hold beta Z)
    const int64_t wgs = beta != 0.0 ? 768 : 1024;
    const int64_t nstripe = std::max<int64_t>(1, std::min<int64_t>(nrb, wgs / (int64_t)ct));
    const dim3 grid((unsigned)round_up(nstripe * ct, kXcds));
    const int ntc = by >= kTile ? 4 : (by + 15) / 16;
#define RCG_TALL_NTC(NTC)                                                                           \
  case NTC:                                                                                         \
    if (beta != 0.0)                                                                                \
      hipLaunchKernelGGL((tall_gemm_mfma_kernel<true, NTC>), grid, dim3(kThreads), 0, st, n, X, ldx, \
                         bx, C, ldc, by, alpha, beta, Z, out, ldo, nrb, (int)ct, nstripe);          \
    else                                                                                            \
      hipLaunchKernelGGL((tall_gemm_mfma_kernel<false, NTC>), grid, dim3(kThreads), 0, st, n, X,    \
                         ldx, bx, C, ldc, by, alpha, beta, Z, out, ldo, nrb, (int)ct, nstripe);     \
    break;
    switch (ntc) {
      RCG_TALL_NTC(1)
      RCG_TALL_NTC(2)
      RCG_TALL_NTC(3)
      RCG_TALL_NTC(4)
      default: throw Error(1, "bad tall product width");
    }
#undef RCG_TALL_NTC
  } else {
    hipLaunchKernelGGL(tall_gemm_kernel<T>, dim3((unsigned)grid, ct), dim3(kThreads), 0, st, n, X,
                       ldx, bx, C, ldc, by, alpha, beta, Z, out, ldo, nrb);
  }
  RCG_HIP(hipGetLastError());
}

size_t colred_ws_elems(int64_t n, int nc) {
  return (size_t)((n + kColRows - 1) / kColRows) * (size_t)std::max(nc, 1);
}

template <typename T>
void launch_col_sumsq(int64_t n, const T* X, int ld, int nc, double* sums, double* ws,
                      size_t ws_elems, hipStream_t st) {
  if (nc <= 0) return;
  const int64_t nblk = (n + kColRows - 1) / kColRows;
  if ((size_t)(nblk * nc) > ws_elems) throw Error(3, "column-reduction workspace too small");
  hipLaunchKernelGGL((col_sumsq_kernel<T, false>), dim3((unsigned)nblk), dim3(kThreads), 0, st, n,
                     X, (const T*)nullptr, ld, nc, (const double*)nullptr, ws);
  RCG_HIP(hipGetLastError());
  hipLaunchKernelGGL((reduce_partials_kernel<double, double>), dim3(reduce_grid(nc)), dim3(kThreads),
                     0, st, ws, nblk, (int64_t)nc, sums);
  RCG_HIP(hipGetLastError());
}

template <typename T>
void launch_residual_sumsq(int64_t n, const T* V, const T* W, int ld, int nc, const double* theta,
                           double* sums, double* ws, size_t ws_elems, hipStream_t st) {
  if (nc <= 0) return;
  const int64_t nblk = (n + kColRows - 1) / kColRows;
  if ((size_t)(nblk * nc) > ws_elems) throw Error(3, "column-reduction workspace too small");
  hipLaunchKernelGGL((col_sumsq_kernel<T, true>), dim3((unsigned)nblk), dim3(kThreads), 0, st, n, V,
                     W, ld, nc, theta, ws);
  RCG_HIP(hipGetLastError());
  hipLaunchKernelGGL((reduce_partials_kernel<double, double>), dim3(reduce_grid(nc)), dim3(kThreads),
                     0, st, ws, nblk, (int64_t)nc, sums);
  RCG_HIP(hipGetLastError());
}

__global__ __launch_bounds__(kThreads) void group_sqrt_kernel(const double* __restrict__ sums,
                                                              int cw, int count,
                                                              double* __restrict__ out) {
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c >= count) return;
  double t = 0.0;
  for (int e = 0; e < cw; ++e) t += sums[cw * c + e];
  out[c] = sqrt(fabs(t));
}

void launch_group_sqrt(const double* sums, int cw, int count, double* out, hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(group_sqrt_kernel, dim3((unsigned)((count + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, st, sums, cw, count, out);
  RCG_HIP(hipGetLastError());
}

template <typename T>
void launch_resid_norm_sumsq(int64_t n, const T* V, const T* W, int ld, int nc,
                             const double* theta, double* sums, double* ws, size_t ws_elems,
                             hipStream_t st) {
  if (nc <= 0) return;
  const int64_t nblk = (n + kColRows - 1) / kColRows;
  if ((size_t)(nblk * 2 * nc) > ws_elems) throw Error(3, "column-reduction workspace too small");
  hipLaunchKernelGGL((resid_norm_kernel<T, const double*>), dim3((unsigned)nblk), dim3(kThreads),
                     0, st, n, V, W, ld, nc, theta, ws);
  RCG_HIP(hipGetLastError());
  hipLaunchKernelGGL((reduce_partials_kernel<double, double>), dim3(reduce_grid(2 * nc)),
                     dim3(kThreads), 0, st, ws, nblk, (int64_t)(2 * nc), sums);
  RCG_HIP(hipGetLastError());
}

void launch_resid_norm_sumsq_args(int64_t n, const double* V, const double* W, int ld, int nc,
                                  const ThetaArgs& theta, double* sums, double* ws,
                                  size_t ws_elems, hipStream_t st) {
  if (nc <= 0) return;
  if (nc > kThetaArgs) throw Error(1, "more Ritz values than the argument block holds");
  const int64_t nblk = (n + kColRows - 1) / kColRows;
  if ((size_t)(nblk * 2 * nc) > ws_elems) throw Error(3, "column-reduction workspace too small");
  hipLaunchKernelGGL((resid_norm_kernel<double, ThetaArgs>), dim3((unsigned)nblk), dim3(kThreads),
                     0, st, n, V, W, ld, nc, theta, ws);
  RCG_HIP(hipGetLastError());
  hipLaunchKernelGGL((reduce_partials_kernel<double, double>), dim3(reduce_grid(2 * nc)),
                     dim3(kThreads), 0, st, ws, nblk, (int64_t)(2 * nc), sums);
  RCG_HIP(hipGetLastError());
}

template <typename T>
void launch_normalize_cols(int64_t n, T* X, int ld, int nc, const double* sums, hipStream_t st) {
  if (n <= 0 || nc <= 0) return;
  hipLaunchKernelGGL(normalize_cols_kernel<T>, dim3(grid_for(n * nc)), dim3(kThreads), 0, st, n, X,
                     ld, nc, sums);
  RCG_HIP(hipGetLastError());
}

template <typename T>
void launch_copy_cols(int64_t n, const T* src, int lds, int s0, T* dst, int ldd, int d0, int cnt,
                      hipStream_t st) {
  if (n <= 0 || cnt <= 0) return;
  hipLaunchKernelGGL(copy_cols_kernel<T>, dim3(grid_for(n * cnt)), dim3(kThreads), 0, st, n, src,
                     lds, s0, dst, ldd, d0, cnt);
  RCG_HIP(hipGetLastError());
}

template <typename T>
void launch_randomize_cols(int64_t n, T* X, int ld, int c0, int cnt, uint64_t seed,
                           uint64_t counter, hipStream_t st) {
  if (n <= 0 || cnt <= 0) return;
  const uint64_t key = [&] {
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }();
  hipLaunchKernelGGL(randomize_cols_kernel<T>, dim3(grid_for(n * cnt)), dim3(kThreads), 0, st, n,
                     X, ld, c0, cnt, key, counter);
  RCG_HIP(hipGetLastError());
}

template <typename T>
void launch_fill_cols(int64_t n, T* X, int ld, int c0, int cnt, T value, hipStream_t st) {
  if (n <= 0 || cnt <= 0) return;
  hipLaunchKernelGGL(fill_cols_kernel<T>, dim3(grid_for(n * cnt)), dim3(kThreads), 0, st, n, X, ld,
                     c0, cnt, value);
  RCG_HIP(hipGetLastError());
}

size_t lz_ws_elems(int64_t n) { return 2 * (size_t)((n + lz_rows(n) - 1) / lz_rows(n)) + 4; }

void launch_lz_dot(int64_t n, const double* w, const double* v, double* out, double* ws,
                   size_t ws_elems, hipStream_t st) {
  const int64_t rpb = lz_rows(n);
  const int64_t nblk = (n + rpb - 1) / rpb;
  if ((size_t)nblk > ws_elems) throw Error(3, "Lanczos workspace too small");
  hipLaunchKernelGGL(lz_dot_kernel, dim3((unsigned)nblk), dim3(kThreads), 0, st, n, w, v, ws, rpb);
  hipLaunchKernelGGL(lz_finalize_kernel, dim3(1), dim3(kThreads), 0, st, ws, nblk, out, 0);
  RCG_HIP(hipGetLastError());
}

void launch_randomize_cplx(int64_t n, double* X, int ld, int c0, int cnt, uint64_t seed,
                           uint64_t counter, hipStream_t st) {
  if (n <= 0 || cnt <= 0) return;
  const uint64_t key = [&] {
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }();
  hipLaunchKernelGGL(randomize_cplx_kernel, dim3(grid_for(n * 2 * cnt)), dim3(kThreads), 0, st, n, X,
                     ld, c0, cnt, key, counter);
  RCG_HIP(hipGetLastError());
}

void launch_pair_sums(double* s, int nc, hipStream_t st) {
  if (nc <= 0) return;
  hipLaunchKernelGGL(pair_sums_kernel, dim3((nc + 255) / 256), dim3(256), 0, st, s, nc);
  RCG_HIP(hipGetLastError());
}

void launch_herm_embed(double* G, int f, int b, hipStream_t st) {
  if (f <= 0 || b <= 0) return;
  hipLaunchKernelGGL(herm_embed_kernel, dim3((f * b + 255) / 256), dim3(256), 0, st, G, f, b);
  RCG_HIP(hipGetLastError());
}

void launch_lz_apply_dot_c(int64_t n, const int* rowptr, const int* col, const double* val,
                           const double* w, double* v, double* alpha_k, double* alpha_im_k,
                           double* ws, size_t ws_elems, hipStream_t st) {
  const int64_t rpb = lz_rows(n);
  const int64_t nblk = (n + rpb - 1) / rpb;
  if ((size_t)(2 * nblk) > ws_elems) throw Error(3, "Lanczos workspace too small");
  hipLaunchKernelGGL(lz_apply_dot_c_kernel, dim3((unsigned)nblk), dim3(kThreads), 0, st, n, rowptr,
                     col, val, w, v, ws, nblk, rpb);
  hipLaunchKernelGGL(lz_finalize_kernel, dim3(1), dim3(kThreads), 0, st, ws, nblk, alpha_k, 0);
  hipLaunchKernelGGL(lz_finalize_kernel, dim3(1), dim3(kThreads), 0, st, ws + nblk, nblk, alpha_im_k, 0);
  RCG_HIP(hipGetLastError());
}



void launch_lz_start(int64_t n, double* v, double* w, double* ws, size_t ws_elems,
                     double* scratch, hipStream_t st) {
  const int64_t rpb = lz_rows(n);
  const int64_t nblk = (n + rpb - 1) / rpb;
  if ((size_t)nblk > ws_elems) throw Error(3, "Lanczos workspace too small");
  hipLaunchKernelGGL(lz_sumsq_kernel, dim3((unsigned)nblk), dim3(kThreads), 0, st, n, v, ws, rpb);
  hipLaunchKernelGGL(lz_finalize_kernel, dim3(1), dim3(kThreads), 0, st, ws, nblk, scratch, 1);
  hipLaunchKernelGGL(lz_init_kernel, dim3(grid_for(n)), dim3(kThreads), 0, st, n, v, w, scratch);
  RCG_HIP(hipGetLastError());
}
void launch_lz_swap(int64_t n, double* v, double* w, const double* beta_prev, hipStream_t st) {
  hipLaunchKernelGGL(lz_swap_kernel, dim3(grid_for(n)), dim3(kThreads), 0, st, n, v, w, beta_prev);
  RCG_HIP(hipGetLastError());
}
void launch_lz_apply_dot(int64_t n, const int* rowptr, const int* col, const double* val,
                         const double* w, double* v, double* alpha_k, double* ws, size_t ws_elems,
                         hipStream_t st) {
  const int64_t rpb = lz_rows(n);
  const int64_t nblk = (n + rpb - 1) / rpb;
  if ((size_t)nblk > ws_elems) throw Error(3, "Lanczos workspace too small");
  hipLaunchKernelGGL(lz_apply_dot_kernel, dim3((unsigned)nblk), dim3(kThreads), 0, st, n, rowptr,
                     col, val, w, v, ws, rpb);
  hipLaunchKernelGGL(lz_finalize_kernel, dim3(1), dim3(kThreads), 0, st, ws, nblk, alpha_k, 0);
  RCG_HIP(hipGetLastError());
}
void launch_lz_axpy_norm(int64_t n, const double* w, double* v, const double* alpha_k,
                         double* beta_k, double* ws, size_t ws_elems, hipStream_t st) {
  const int64_t rpb = lz_rows(n);
  const int64_t nblk = (n + rpb - 1) / rpb;
  if ((size_t)nblk > ws_elems) throw Error(3, "Lanczos workspace too small");
  hipLaunchKernelGGL(lz_axpy_norm_kernel, dim3((unsigned)nblk), dim3(kThreads), 0, st, n, w, v,
                     alpha_k, ws, rpb);
  hipLaunchKernelGGL(lz_finalize_kernel, dim3(1), dim3(kThreads), 0, st, ws, nblk, beta_k, 1);
  RCG_HIP(hipGetLastError());
}

void launch_lz_step(int64_t n, const int* rowptr, const int* col, const double* val,
                    const double* v, const double* w, double* vn, double* wn,
                    const double* beta_prev, double* alpha_k, double* beta_k, unsigned* cnt,
                    double* ws, size_t ws_elems, hipStream_t st) {
  const int64_t rpb = lz_rows(n);
  const int64_t nblk = (n + rpb - 1) / rpb;
  if ((size_t)nblk > ws_elems) throw Error(3, "Lanczos workspace too small");
  hipLaunchKernelGGL(lz_step_kernel, dim3((unsigned)nblk), dim3(kThreads), 0, st, n, rowptr, col,
                     val, v, w, vn, wn, beta_prev, ws, cnt, alpha_k, rpb);
  hipLaunchKernelGGL(lz_axpy_beta_kernel, dim3((unsigned)nblk), dim3(kThreads), 0, st, n, wn, vn,
                     alpha_k, ws, cnt, beta_k, rpb);
  RCG_HIP(hipGetLastError());
}

// ------------------------------------------------------- instantiations
// ------------------------------------------ layouts for caller operators
// A caller-supplied operator works on column-major blocks (element (i, c) at
// C[c*ldc + i], the reference's AmatrixClassLMCuda layout); the solver's
// blocks are row-interleaved (element (i, c) at X[i*ldx + c]).  E is one
// block entry: double, or double2 for a complex entry (re, im).  32 x 32
// tiles through LDS keep both sides coalesced.
template <typename E>
struct EntryOps;
template <>
struct EntryOps<double> {
  __device__ static double comb(double p0, double p1, double p2, double ax, double x, double z) {
    return p2 * fma(p1, fma(p0, ax, x), -z);
  }
  __device__ static double zero() { return 0.0; }
};
template <>
struct EntryOps<double2> {
  __device__ static double2 comb(double p0, double p1, double p2, double2 ax, double2 x,
                                 double2 z) {
    return make_double2(p2 * fma(p1, fma(p0, ax.x, x.x), -z.x), p2 * fma(p1, fma(p0, ax.y, x.y), -z.y));
  }
  __device__ static double2 zero() { return make_double2(0.0, 0.0); }
};

constexpr int kLt = 32;  // layout tile

// TO_COLS: rows layout -> column-major; else the reverse.
template <typename E, bool TO_COLS>
__global__ __launch_bounds__(kThreads) void layout_kernel(int64_t n, int cols, E* R, int64_t ldr,
                                                          E* Cm, int64_t ldc) {
  __shared__ E tile[kLt][kLt + 1];
  const int64_t i0 = (int64_t)blockIdx.x * kLt;
  const int c0 = blockIdx.y * kLt;
  const int tx = threadIdx.x % kLt, ty = threadIdx.x / kLt;
  for (int r = ty; r < kLt; r += kThreads / kLt) {
    if (TO_COLS) {  // tile[row][col] from the rows layout
      const int64_t i = i0 + r;
      const int c = c0 + tx;
      if (i < n && c < cols) tile[r][tx] = R[i * ldr + c];
    } else {  // tile[col][row] from the column-major layout
      const int c = c0 + r;
      const int64_t i = i0 + tx;
      if (i < n && c < cols) tile[r][tx] = Cm[(int64_t)c * ldc + i];
    }
  }
  __syncthreads();
  for (int r = ty; r < kLt; r += kThreads / kLt) {
    if (TO_COLS) {
      const int c = c0 + r;
      const int64_t i = i0 + tx;
      if (i < n && c < cols) Cm[(int64_t)c * ldc + i] = tile[tx][r];
    } else {
      const int64_t i = i0 + r;
      const int c = c0 + tx;
      if (i < n && c < cols) R[i * ldr + c] = tile[tx][r];
    }
  }
}

// out = p2 (p1 (p0 AX + X) - Z) in the rows layout, AX column-major (Z may
// be null: no Z term; out may alias Z): the Chebyshev step around a caller
// operator's product, the same expression as the fused CSR step's epilogue.
template <typename E>
__global__ __launch_bounds__(kThreads) void cheb_combine_kernel(int64_t n, int cols,
                                                                const E* __restrict__ AX,
                                                                int64_t ldc, const E* X,
                                                                const E* Z, E* out, int64_t ldr,
                                                                double p0, double p1, double p2) {
  __shared__ E tile[kLt][kLt + 1];
  const int64_t i0 = (int64_t)blockIdx.x * kLt;
  const int c0 = blockIdx.y * kLt;
  const int tx = threadIdx.x % kLt, ty = threadIdx.x / kLt;
  for (int r = ty; r < kLt; r += kThreads / kLt) {
    const int c = c0 + r;
    const int64_t i = i0 + tx;
    if (i < n && c < cols) tile[r][tx] = AX[(int64_t)c * ldc + i];
  }
  __syncthreads();
  for (int r = ty; r < kLt; r += kThreads / kLt) {
    const int64_t i = i0 + r;
    const int c = c0 + tx;
    if (i < n && c < cols) {
      const int64_t o = i * ldr + c;
      const E z = Z ? Z[o] : EntryOps<E>::zero();
      out[o] = EntryOps<E>::comb(p0, p1, p2, tile[tx][r], X[o], z);
    }
  }
}

void launch_layout(int64_t n, int cols, int cw, double* rows, int64_t ldr_real, double* colmaj,
                   int64_t ldc, bool to_cols, hipStream_t st) {
  if (n <= 0 || cols <= 0) return;
  if (cw != 1 && cw != 2) throw Error(1, "entries are real or complex");
  if (ldr_real % cw != 0) throw Error(1, "row stride is not a whole number of entries");
  const dim3 grid((unsigned)((n + kLt - 1) / kLt), (unsigned)((cols + kLt - 1) / kLt));
  const int64_t ldr = ldr_real / cw;
  if (cw == 1) {
    if (to_cols)
      hipLaunchKernelGGL((layout_kernel<double, true>), grid, dim3(kThreads), 0, st, n, cols, rows,
                         ldr, colmaj, ldc);
    else
      hipLaunchKernelGGL((layout_kernel<double, false>), grid, dim3(kThreads), 0, st, n, cols, rows,
                         ldr, colmaj, ldc);
  } else {
    double2* r2 = reinterpret_cast<double2*>(rows);
    double2* c2 = reinterpret_cast<double2*>(colmaj);
    if (to_cols)
      hipLaunchKernelGGL((layout_kernel<double2, true>), grid, dim3(kThreads), 0, st, n, cols, r2,
                         ldr, c2, ldc);
    else
      hipLaunchKernelGGL((layout_kernel<double2, false>), grid, dim3(kThreads), 0, st, n, cols, r2,
                         ldr, c2, ldc);
  }
  RCG_HIP(hipGetLastError());
}

void launch_cheb_combine(int64_t n, int cols, int cw, const double* AX, int64_t ldc,
                         const double* X, const double* Z, double* out, int64_t ldr_real,
                         double p0, double p1, double p2, hipStream_t st) {
  if (n <= 0 || cols <= 0) return;
  if (cw != 1 && cw != 2) throw Error(1, "entries are real or complex");
  if (ldr_real % cw != 0) throw Error(1, "row stride is not a whole number of entries");
  const dim3 grid((unsigned)((n + kLt - 1) / kLt), (unsigned)((cols + kLt - 1) / kLt));
  const int64_t ldr = ldr_real / cw;
  if (cw == 1) {
    hipLaunchKernelGGL(cheb_combine_kernel<double>, grid, dim3(kThreads), 0, st, n, cols, AX, ldc,
                       X, Z, out, ldr, p0, p1, p2);
  } else {
    hipLaunchKernelGGL(cheb_combine_kernel<double2>, grid, dim3(kThreads), 0, st, n, cols,
                       reinterpret_cast<const double2*>(AX), ldc,
                       reinterpret_cast<const double2*>(X), reinterpret_cast<const double2*>(Z),
                       reinterpret_cast<double2*>(out), ldr, p0, p1, p2);
  }
  RCG_HIP(hipGetLastError());
}

// partial sums of Re and Im of sum conj(w) v over interleaved complex vectors
__global__ __launch_bounds__(kThreads) void lz_dot_c_kernel(int64_t n, const double* w,
                                                            const double* v, double* P,
                                                            int64_t nblk, int64_t rpb) {
  __shared__ double red[kThreads];
  const int64_t i0 = blockIdx.x * rpb;
  const int64_t i1 = min(n, i0 + rpb);
  double sr = 0.0, si = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
    const double wr = w[2 * i], wi = w[2 * i + 1], vr = v[2 * i], vi = v[2 * i + 1];
    sr = fma(wr, vr, fma(wi, vi, sr));
    si = fma(wr, vi, fma(-wi, vr, si));
  }
  sr = block_sum(sr, red);
  si = block_sum(si, red);
  if (threadIdx.x == 0) {
    P[blockIdx.x] = sr;
    P[nblk + blockIdx.x] = si;
  }
}

void launch_lz_dot_c(int64_t n, const double* w, const double* v, double* alpha_k,
                     double* alpha_im_k, double* ws, size_t ws_elems, hipStream_t st) {
  const int64_t rpb = lz_rows(n);
  const int64_t nblk = (n + rpb - 1) / rpb;
  if ((size_t)(2 * nblk) > ws_elems) throw Error(3, "Lanczos workspace too small");
  hipLaunchKernelGGL(lz_dot_c_kernel, dim3((unsigned)nblk), dim3(kThreads), 0, st, n, w, v, ws,
                     nblk, rpb);
  hipLaunchKernelGGL(lz_finalize_kernel, dim3(1), dim3(kThreads), 0, st, ws, nblk, alpha_k, 0);
  hipLaunchKernelGGL(lz_finalize_kernel, dim3(1), dim3(kThreads), 0, st, ws + nblk, nblk,
                     alpha_im_k, 0);
  RCG_HIP(hipGetLastError());
}

#define RCG_INST(T)                                                                              \
  template void launch_csr_apply<T>(int64_t, const int*, const int*, const T*, const T*, T*, int, \
                                    int, T, T, hipStream_t, const CsrSched*, bool);               \
  template void launch_cheb_step<T>(int64_t, const int*, const int*, const T*, const T*,          \
                                    const T*, T*, int, int, T, T, T, bool, hipStream_t, int,      \
                                    const CsrSched*, bool);                                       \
  template size_t gram_ws_elems<T>(int64_t, int, int);                                           \
  template void launch_gram<T>(int64_t, const T*, int, int, const T*, int, int, T*, T*, size_t,   \
                               hipStream_t, bool);                                                \
  template void launch_tall_gemm<T>(int64_t, const T*, int, int, const T*, int, int, T, T,        \
                                    const T*, T*, int, hipStream_t);                              \
  template void launch_col_sumsq<T>(int64_t, const T*, int, int, double*, double*, size_t,        \
                                    hipStream_t);                                                 \
  template void launch_residual_sumsq<T>(int64_t, const T*, const T*, int, int, const double*,    \
                                         double*, double*, size_t, hipStream_t);                  \
  template void launch_normalize_cols<T>(int64_t, T*, int, int, const double*, hipStream_t);      \
  template void launch_resid_norm_sumsq<T>(int64_t, const T*, const T*, int, int, const double*,  \
                                           double*, double*, size_t, hipStream_t);                \
  template void launch_copy_cols<T>(int64_t, const T*, int, int, T*, int, int, int, hipStream_t); \
  template void launch_randomize_cols<T>(int64_t, T*, int, int, int, uint64_t, uint64_t,          \
                                         hipStream_t);                                            \
  template void launch_fill_cols<T>(int64_t, T*, int, int, int, T, hipStream_t);                 \
  template void launch_dense_step<T>(int64_t, const T*, int64_t, const T*, const T*, T*, int, int, \
                                     T, T, T, int, const T*, const T*, hipStream_t);

RCG_INST(double)
RCG_INST(float)

}  // namespace rcg
