// rcg_csr_rows.hpp -- device code shared by the CSR filter kernels of
// rcg_kernels.hip (one operator per launch) and rcg_batch.hip (many): launch
// constants, the XCD remap, 16-byte row-segment access with non-temporal
// single-use streams, and csr_rows, the fused SpMM / Chebyshev step over a
// row block.  Each translation unit keeps its own copy (anonymous namespace).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "rcg_internal.hpp"

namespace rcg {
namespace {

constexpr int kXcds = 8;
constexpr int kThreads = 256;

// Logical block for physical block `bid` of `nb` (nb a multiple of 8).  The
// hardware deals workgroups round-robin to the 8 XCDs; this mapping gives
// XCD x the logical blocks [x*nb/8, (x+1)*nb/8), i.e. contiguous rows.
__device__ __forceinline__ int64_t xcd_remap(int64_t bid, int64_t nb) {
#ifdef RCG_NO_XCD_REMAP
  return bid;
#else
  const int64_t per = nb / kXcds;
  return (bid % kXcds) * per + bid / kXcds;
#endif
}

// Non-temporal single-use streams: 3% faster FP64 and FP32 filter steps at
// 128^3 (the gathered rows keep more of L2); RCG_TEMPORAL_STREAMS turns them off.
#ifdef RCG_TEMPORAL_STREAMS
constexpr bool kNtStreams = false;
#else
constexpr bool kNtStreams = true;
#endif


// A single-use scalar (the CSR arrays, read once per step), non-temporal
// with the streams above (FP64 step at 128^3: 944 -> 943 us).
template <typename T>
__device__ __forceinline__ T ld_once(const T* p) {
  if constexpr (kNtStreams) return __builtin_nontemporal_load(p);
  else return *p;
}

// 16-byte vector access to a row segment.
template <typename T>
struct V16;
// ld/st go through the caches normally; ld_stream/st_stream mark single-use
// data (the X_{j-2} read and the X_j write of a step) non-temporal, so L2
// keeps the gathered X_{j-1} rows.
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
template <>
struct V16<double> {
  static constexpr int W = 2;
  __device__ static void ld(const double* p, double (&o)[2]) {
    const dvec2 t = *reinterpret_cast<const dvec2*>(p);
    o[0] = t.x;
    o[1] = t.y;
  }
  __device__ static void st(double* p, const double (&o)[2]) {
    *reinterpret_cast<dvec2*>(p) = dvec2{o[0], o[1]};
  }
  __device__ static void ld_stream(const double* p, double (&o)[2]) {
    if constexpr (kNtStreams) {
      const dvec2 t = __builtin_nontemporal_load(reinterpret_cast<const dvec2*>(p));
      o[0] = t.x;
      o[1] = t.y;
    } else {
      ld(p, o);
    }
  }
  __device__ static void st_stream(double* p, const double (&o)[2]) {
    if constexpr (kNtStreams) __builtin_nontemporal_store(dvec2{o[0], o[1]}, reinterpret_cast<dvec2*>(p));
    else st(p, o);
  }
};
template <>
struct V16<float> {
  static constexpr int W = 4;
  __device__ static void ld(const float* p, float (&o)[4]) {
    const fvec4 t = *reinterpret_cast<const fvec4*>(p);
    o[0] = t.x;
    o[1] = t.y;
    o[2] = t.z;
    o[3] = t.w;
  }
  __device__ static void st(float* p, const float (&o)[4]) {
    *reinterpret_cast<fvec4*>(p) = fvec4{o[0], o[1], o[2], o[3]};
  }
  __device__ static void ld_stream(const float* p, float (&o)[4]) {
    if constexpr (kNtStreams) {
      const fvec4 t = __builtin_nontemporal_load(reinterpret_cast<const fvec4*>(p));
      o[0] = t.x;
      o[1] = t.y;
      o[2] = t.z;
      o[3] = t.w;
    } else {
      ld(p, o);
    }
  }
  __device__ static void st_stream(float* p, const float (&o)[4]) {
    if constexpr (kNtStreams) __builtin_nontemporal_store(fvec4{o[0], o[1], o[2], o[3]}, reinterpret_cast<fvec4*>(p));
    else st(p, o);
  }
};

// ------------------------------------------------------------ CSR x block
// kChebSrc: kCheb plus a per-column multiple of a source block,
//   out = p2 * (p1*(p0*A*X + X) - Z + kap[c] * R),
// the correction recurrence of the mixed-precision filter.
// kApplyStart (FP64): W = A V, whose entries go straight into the mixed
// filter's start (mixed_start_elem) instead of being stored.
// kChebSrcEnd (FP32): kChebSrc whose output x goes straight into the mixed
// filter's recombination, V <- s_end * (um[c] V + x) (V FP64), not stored.
enum CsrMode {
  kApply = 0,
  kApplyBeta = 1,
  kChebFirst = 2,
  kCheb = 3,
  kChebSrc = 4,
  kApplyStart = 5,
  kChebSrcEnd = 6
};

// The mixed filter's start for one element, v = V[i][c] and w = (A V)[i][c]:
// a Ritz column (flag != 0) gets R = w - theta v, E0 = 0, E1 = e1c R; a plain
// column R = 0, E0 = v, E1 = g2 (g1 w + v).  Shared by the fused apply and
// the stand-alone start kernel, so that both round alike.
__device__ __forceinline__ void mixed_start_elem(double v, double w, double theta, double flag,
                                                 double e1c, double g1, double g2, float& e0,
                                                 float& e1, float& r) {
  if (flag != 0.0) {
    const double d = fma(-theta, v, w);
    r = float(d);
    e0 = 0.0f;
    e1 = float(e1c * d);
  } else {
    r = 0.0f;
    e0 = float(v);
    e1 = float(g2 * fma(g1, w, v));
  }
}

// Rows per workgroup block (cap) and nonzeros staged in LDS per block.
constexpr int kMaxRpb = 512;
constexpr int kStageNz = 1024;
// Nonzeros of a row whose gathers are issued together before their FMAs, by
// kernel (a template argument of csr_rows; the FMAs run in the same order for
// every chunk size).  Measured at 128^3 x 64 and on 32 served operators:
// single-operator FP64 step 4 (830 us; 3: 841, 2: 856, 8: +3%), real FP32
// correction step 2 (512 us; 1: 530, 3: 523, 4: 537), batched FP64 step 3
// (8.39 ms per 24-step call; 2: 8.61, 4: 8.88, 8: 11.5); at 1024^2 x 96 complex
// columns,
// complex128 step 3 (1043 us; 2: 1079, 4: 1066) and complex64 step 3 (708 us;
// 2: 720, 4: 770).  More gathers in flight cost the address unit more than
// the trips to memory they save.
#ifndef RCG_CSR_CHUNK
#define RCG_CSR_CHUNK 4
#endif
#ifndef RCG_CSR_CHUNK_F32
#define RCG_CSR_CHUNK_F32 2
#endif
#ifndef RCG_CSR_CHUNK_C64
#define RCG_CSR_CHUNK_C64 3
#endif
#ifndef RCG_CSR_CHUNK_C128
#define RCG_CSR_CHUNK_C128 3
#endif
constexpr int kChunk = RCG_CSR_CHUNK;
// the single-operator step's chunk by element type
template <typename T, bool CX>
constexpr int csr_chunk() {
  if (CX) return std::is_same<T, float>::value ? RCG_CSR_CHUNK_C64 : RCG_CSR_CHUNK_C128;
  return std::is_same<T, float>::value ? RCG_CSR_CHUNK_F32 : kChunk;
}

// G lanes own one row; each lane handles NV 16-byte column vectors at
// cv = g + k*G.  256/G rows are in flight per workgroup.  The block's row
// pointers (and, when they fit, its column indices and values) are staged in
// LDS by one coalesced load, and a row's own-row operands are loaded before
// its gathers, so that computing a row costs a single trip to memory.
// CX: complex operator and block (interleaved re, im): each 16-byte vector
// holds W/2 complex entries and the gather multiplies complex values.
// tail (1..W): entries of the last vector (cv = ncv - 1) that are columns of
// the call; the rest of that vector is read but never written.
// Rows [i0, i1) of one step (the body shared by the single-operator and the
// batched kernels); s_rp / s_col / s_val are the workgroup's LDS stages.
template <typename T, int G, int NV, int MODE, int RPI, bool CX, int CH>
__device__ __forceinline__ void csr_rows(
    int64_t n, const int* __restrict__ rowptr, const int* __restrict__ col,
    const T* __restrict__ val, const T* __restrict__ X, const T* Z, T* out, int ld, int ncv,
    int tail, T p0, T p1, T p2, int64_t i0, int64_t i1, const T* __restrict__ R,
    const T* __restrict__ kap, int* s_rp, int* s_col, T* s_val, const MixedArgs& ms) {
  using VT = V16<T>;
  constexpr int W = VT::W;
  constexpr int RS = kThreads / G;
  static_assert(MODE != kApplyStart || std::is_same<T, double>::value, "the start is FP64");
  static_assert(MODE != kChebSrcEnd || std::is_same<T, float>::value, "the end follows FP32 steps");
  constexpr bool kSrc = MODE == kChebSrc || MODE == kChebSrcEnd;
  const int g = threadIdx.x % G;
  const int r = threadIdx.x / G;
  // kApplyStart: this lane's columns' (theta, flag), the same for every row
  double th[MODE == kApplyStart ? NV : 1][W], fl[MODE == kApplyStart ? NV : 1][W];
  if constexpr (MODE == kApplyStart) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const int c = (g + k * G) * W + w;
        th[k][w] = c < ms.nc ? ms.cp[c] : 0.0;
        fl[k][w] = c < ms.nc ? ms.cp[ms.nc + c] : 0.0;
      }
  }
  const int nr = (int)(i1 - i0);
  for (int t = threadIdx.x; t <= nr; t += kThreads) s_rp[t] = ld_once(rowptr + i0 + t);
  __syncthreads();
  const int base = s_rp[0];
  const int nz = s_rp[nr] - base;
  const bool staged = nz <= kStageNz;
  if (staged)
    for (int t = threadIdx.x; t < nz; t += kThreads) {
      s_col[t] = ld_once(col + base + t);
      if constexpr (CX) {
        s_val[2 * t] = ld_once(val + 2 * (int64_t)(base + t));
        s_val[2 * t + 1] = ld_once(val + 2 * (int64_t)(base + t) + 1);
      } else {
        s_val[t] = ld_once(val + base + t);
      }
    }
  __syncthreads();
  // kChebSrc: this lane's columns' kap (and, ending, um), the same for every
  // row (loaded once, not with each row's operands)
  T kv[kSrc ? NV : 1][W];
  double uv[MODE == kChebSrcEnd ? NV : 1][W];
  if constexpr (kSrc) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const int cv = g + k * G;
        kv[k][w] = cv < ncv ? kap[cv * W + w] : T(0);
        if constexpr (MODE == kChebSrcEnd) uv[k][w] = cv < ncv ? ms.um[cv * W + w] : 0.0;
      }
  }
  // RPI rows per lane iteration: rows li and li + RS, ... so that each wave
  // has RPI rows' gathers in flight at once.  When G does not divide the
  // workgroup, the last kThreads % G lanes own no row.
  const int rfirst = (kThreads % G == 0 || r < RS) ? r : nr;
  for (int li0 = rfirst; li0 < nr; li0 += RS * RPI) {
    int pb[RPI], pe[RPI];
    bool live_row[RPI];
#pragma unroll
    for (int q = 0; q < RPI; ++q) {
      const int li = li0 + q * RS;
      live_row[q] = li < nr;
      pb[q] = live_row[q] ? s_rp[li] : 0;
      pe[q] = live_row[q] ? s_rp[li + 1] : 0;
    }
    // own-row operands first: they do not depend on the gathers
    T xs[RPI][NV][W], zs[RPI][NV][W];
    T rs[kSrc ? RPI : 1][NV][W];
#pragma unroll
    for (int q = 0; q < RPI; ++q) {
      if (!live_row[q]) continue;
      const int64_t rowoff = (i0 + li0 + q * RS) * (int64_t)ld;
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        const int cv = g + k * G;
        if (cv < ncv) {
          if constexpr (MODE == kChebFirst || MODE == kCheb || kSrc || MODE == kApplyStart)
            VT::ld(X + rowoff + cv * W, xs[q][k]);
          if constexpr (MODE == kCheb || kSrc) VT::ld_stream(Z + rowoff + cv * W, zs[q][k]);
          if constexpr (kSrc) VT::ld_stream(R + rowoff + cv * W, rs[q][k]);
          if constexpr (MODE == kApplyBeta) VT::ld(out + rowoff + cv * W, zs[q][k]);
        }
      }
    }
    T acc[RPI][NV][W];
#pragma unroll
    for (int q = 0; q < RPI; ++q)
#pragma unroll
      for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int w = 0; w < W; ++w) acc[q][k][w] = T(0);
    int maxlen = 0;
#pragma unroll
    for (int q = 0; q < RPI; ++q) maxlen = max(maxlen, pe[q] - pb[q]);
    // chunks of CH nonzeros per row: every gather of a chunk (over the
    // RPI rows) is issued before its FMAs
    for (int c = 0; c < maxlen; c += CH) {
      int jj[RPI][CH];
      T aa[RPI][CH], ai[RPI][CH];
#pragma unroll
      for (int q = 0; q < RPI; ++q)
#pragma unroll
        for (int e = 0; e < CH; ++e) {
          const int p = pb[q] + c + e;
          const bool live = p < pe[q];
          if constexpr (CX) {
            if (staged) {
              jj[q][e] = live ? s_col[p - base] : 0;
              aa[q][e] = live ? s_val[2 * (p - base)] : T(0);
              ai[q][e] = live ? s_val[2 * (p - base) + 1] : T(0);
            } else {
              jj[q][e] = live ? col[p] : 0;
              aa[q][e] = live ? val[2 * (int64_t)p] : T(0);
              ai[q][e] = live ? val[2 * (int64_t)p + 1] : T(0);
            }
          } else {
            if (staged) {
              jj[q][e] = live ? s_col[p - base] : 0;
              aa[q][e] = live ? s_val[p - base] : T(0);
            } else {
              jj[q][e] = live ? col[p] : 0;
              aa[q][e] = live ? val[p] : T(0);
            }
          }
        }
      T u[RPI][CH][NV][W];
#pragma unroll
      for (int q = 0; q < RPI; ++q)
#pragma unroll
        for (int e = 0; e < CH; ++e) {
          if (pb[q] + c + e < pe[q]) {
            const T* xr = X + (int64_t)jj[q][e] * ld;
#pragma unroll
            for (int k = 0; k < NV; ++k) {
              const int cv = g + k * G;
              if (cv < ncv) VT::ld(xr + cv * W, u[q][e][k]);
            }
          }
        }
#pragma unroll
      for (int q = 0; q < RPI; ++q)
#pragma unroll
        for (int e = 0; e < CH; ++e) {
          if (pb[q] + c + e < pe[q]) {
#pragma unroll
            for (int k = 0; k < NV; ++k) {
              if constexpr (CX) {
#pragma unroll
                for (int w = 0; w < W; w += 2) {
                  acc[q][k][w] = fma(aa[q][e], u[q][e][k][w], acc[q][k][w]);
                  acc[q][k][w] = fma(-ai[q][e], u[q][e][k][w + 1], acc[q][k][w]);
                  acc[q][k][w + 1] = fma(aa[q][e], u[q][e][k][w + 1], acc[q][k][w + 1]);
                  acc[q][k][w + 1] = fma(ai[q][e], u[q][e][k][w], acc[q][k][w + 1]);
                }
              } else {
#pragma unroll
                for (int w = 0; w < W; ++w) acc[q][k][w] = fma(aa[q][e], u[q][e][k][w], acc[q][k][w]);
              }
            }
          }
        }
    }
#pragma unroll
    for (int q = 0; q < RPI; ++q) {
      if (!live_row[q]) continue;
      const int64_t rowoff = (i0 + li0 + q * RS) * (int64_t)ld;
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        const int cv = g + k * G;
        if (cv >= ncv) continue;
        if constexpr (MODE == kApplyStart) {
          float e0[W], e1[W], rr[W];
#pragma unroll
          for (int w = 0; w < W; ++w)
            mixed_start_elem(xs[q][k][w], p0 * acc[q][k][w], th[k][w], fl[k][w], ms.e1c, ms.g1,
                             ms.g2, e0[w], e1[w], rr[w]);
          const int64_t e = rowoff + cv * W;
          if (tail < W && cv == ncv - 1) {
#pragma unroll
            for (int w = 0; w < W; ++w)
              if (w < tail) {
                ms.e0[e + w] = e0[w];
                ms.e1[e + w] = e1[w];
                ms.r[e + w] = rr[w];
              }
          } else {
            typedef float fvec2 __attribute__((ext_vector_type(2)));
            *reinterpret_cast<fvec2*>(ms.e0 + e) = fvec2{e0[0], e0[1]};
            *reinterpret_cast<fvec2*>(ms.e1 + e) = fvec2{e1[0], e1[1]};
            *reinterpret_cast<fvec2*>(ms.r + e) = fvec2{rr[0], rr[1]};
          }
          continue;
        }
        T o[W];
        if constexpr (MODE == kApply) {
#pragma unroll
          for (int w = 0; w < W; ++w) o[w] = p0 * acc[q][k][w];
        } else if constexpr (MODE == kApplyBeta) {
#pragma unroll
          for (int w = 0; w < W; ++w) o[w] = fma(p0, acc[q][k][w], p1 * zs[q][k][w]);
        } else if constexpr (MODE == kChebFirst) {
#pragma unroll
          for (int w = 0; w < W; ++w) o[w] = p2 * (p1 * fma(p0, acc[q][k][w], xs[q][k][w]));
        } else if constexpr (kSrc) {
#pragma unroll
          for (int w = 0; w < W; ++w)
            o[w] = p2 * fma(kv[k][w], rs[q][k][w],
                            fma(p1, fma(p0, acc[q][k][w], xs[q][k][w]), -zs[q][k][w]));
        } else {
#pragma unroll
          for (int w = 0; w < W; ++w)
            o[w] = p2 * fma(p1, fma(p0, acc[q][k][w], xs[q][k][w]), -zs[q][k][w]);
        }
        if constexpr (MODE == kChebSrcEnd) {
          // V <- s_end (um V + x) on this vector's columns (FP64)
          double* vp = ms.v_end + rowoff + cv * W;
          if (tail < W && cv == ncv - 1) {
#pragma unroll
            for (int w = 0; w < W; ++w)
              if (w < tail) vp[w] = ms.s_end * fma(uv[k][w], vp[w], double(o[w]));
          } else {
            // W = 4 floats: four doubles, 32-byte aligned (ld is a multiple of 4)
            const dvec2 v0 = reinterpret_cast<const dvec2*>(vp)[0];
            const dvec2 v1 = reinterpret_cast<const dvec2*>(vp)[1];
            const double vv[4] = {v0.x, v0.y, v1.x, v1.y};
            double r[4];
#pragma unroll
            for (int w = 0; w < 4; ++w) r[w] = ms.s_end * fma(uv[k][w], vv[w], double(o[w]));
            reinterpret_cast<dvec2*>(vp)[0] = dvec2{r[0], r[1]};
            reinterpret_cast<dvec2*>(vp)[1] = dvec2{r[2], r[3]};
          }
        } else if (tail < W && cv == ncv - 1) {
          // the last vector holds columns past the caller's count: only the
          // first `tail` entries are stored
#pragma unroll
          for (int w = 0; w < W; ++w)
            if (w < tail) out[rowoff + cv * W + w] = o[w];
        } else {
          VT::st_stream(out + rowoff + cv * W, o);
        }
      }
    }
  }
}

}  // namespace
}  // namespace rcg
