// cg.hip -- (Jacobi-)preconditioned conjugate gradients on gfx950.
//
// Plain path: the reference's cg_solve_body (LinearAlgebra/IterativeLinearSolver/
// cg.f90:101-238) with all its stopping criteria (cg.f90:279-407), in FP64.
// Three kernels per iteration, each one HBM pass:
//   spmv : q = A p ; block partials of <p,q>, ||p||^2
//   upd  : alpha = rz/<p,q> ; r -= alpha q ; partials <r,z>, ||r||^2
//   dir  : stopping test, beta ; x += alpha p ; p = z + beta p   (z = invD r, never stored)
// Scalars never leave the device.  A kernel reduces the block partials of its
// producer at entry, every block in the same fixed order, so all blocks agree
// bit for bit and no grid-wide fence or atomic is needed (visibility comes from
// the kernel boundary).  The scalar state is double-buffered by iteration
// parity: iteration i reads S[i&1] and its last kernel writes S[(i+1)&1], so no
// block ever reads a field another block of the same launch writes.  Once the
// state says `done`, every kernel returns at entry; the host launches
// hipGraph-captured chunks of iterations and polls a small result block.
//
// The SpMV runs on the sliced diagonal/ELL copy of the CSR matrix (one
// wavefront per 64-row slice, column offsets of diagonal slices are
// wave-uniform, so the matrix stream carries no indices), with each XCD given a
// contiguous range of slices so that the +-plane neighbours of p stay in its L2.
//
// FP32-refined path: CG iterated in FP32 (FP32 matrix copy, FP32 arithmetic and
// accumulation) with reliable updates (residual replacement, van der Vorst & Ye,
// SIAM J. Sci. Comput. 22, 2000): when the FP32 recurrence residual has fallen
// by a factor delta since the last update, the FP32 solution increment is
// added into the FP64 solution, the residual r = b - A x is recomputed in FP64
// on the FP64 matrix and replaces the FP32 residual, and the stopping test is
// made on that FP64 residual.  The search direction is kept, so the Krylov
// convergence rate of plain CG is preserved.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <tuple>
#include <type_traits>

#include "femhip_internal.h"

namespace femhip {
namespace {

#ifndef FEMHIP_CG_THREADS
#define FEMHIP_CG_THREADS 512
#endif
constexpr int kThreads = FEMHIP_CG_THREADS;
constexpr int kWaves = kThreads / 64;
// Elements per thread a vector kernel loads before it reduces its producer's
// block partials, so the two latencies overlap (the grids give about 4
// elements per thread at 128^3).
constexpr int kPrefetch = 4;
// Reliable-update factor of the FP32 iterations: an update when the scaled
// recurrence residual fell by this factor (measured at 128^3: 0.02 gives 263
// inner + 7 updates, 0.1 gives 259 + 11, 0.005 gives 276 + 5).
constexpr double kDefaultDelta = 0.02;

struct CgState {
  double rz, rnrm_m, alpha, beta;
  double rhs_test, rhs_extra, est, pap, pp;
  double bnorm, r0norm, rnorm;
  double rtol, atol;
  double rmax, delta;          // reliable updates
  double ru_rz, ru_rmax, ru_rnorm, ru_beta;
  double rscale, ru_rscale;    // ||S r|| / ||r|| at the last update (scaled inner residual)
  int iter, done, converged, maxit, crit, ru, updates, ru_conv, ru_done, first;
};

struct CgResult {
  double est, rhs_test, rnorm, bnorm;
  int iter, done, converged, updates;
  double prog;                 // latest quantity the stopping test compares with rhs_test ...
  int prog_it, pad;            // ... and its iteration (for the host's chunk planning)
};

struct SlicedDev {
  int32_t ns, n;
  const int64_t *base;
  const int32_t *meta, *aux, *offs, *ecol;
  int32_t sym, kplus;          // symmetric-diagonal mode
  int64_t stride;
  const int32_t *soffs;
  int32_t order;               // traversal order of wave_slices
  const int32_t *srow;         // rows of the lanes of permuted slices
  int32_t pstride;             // order 2: slices per step of a wave (about one plane)
  int32_t trip;                // symmetric 27-point offsets in column triplets (sym27_row)
  int32_t unit;                // symmetric layout: the diagonal is 1 and is not read
  int32_t alt;                 // CG SpMVs: odd iterations sweep the slices in reverse
};

// Row of lane `lane` of slice s (meta bit 28: the slice's rows are a sorted
// window's, listed in srow; a lane without a row gets n).  The symmetric
// layout has no per-slice metadata and formats without sorted windows no row
// list: their rows are in order.
__device__ __forceinline__ int slice_row_index(const SlicedDev &m, int s, int lane) {
  if (m.sym || m.srow == nullptr) return s * kSlice + lane;
  return ((m.meta[s] >> 28) & 1) ? m.srow[s * kSlice + lane] : s * kSlice + lane;
}

__device__ __forceinline__ bool is_delta(int c) {
  return c == FEMHIP_DELTA_RHS || c == FEMHIP_DELTA_DELTA || c >= FEMHIP_DELTA_RHS_AND_RES_RES;
}
__device__ __forceinline__ bool is_delta_delta(int c) {
  return c == FEMHIP_DELTA_DELTA || c == FEMHIP_DELTA_DELTA_AND_RES_RES || c == FEMHIP_DELTA_DELTA_AND_RES_RHS;
}

// Sum NV values over the block; every thread gets the result.
template <int NV>
__device__ __forceinline__ void block_sum_all(double (&v)[NV]) {
  __shared__ double sm[NV][kWaves];
  __shared__ double out[NV];
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < NV; ++k) v[k] += __shfl_down(v[k], off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < NV; ++k) sm[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int i = 0; i < kWaves; ++i) s += sm[k][i];
      out[k] = s;
    }
  __syncthreads();
  for (int k = 0; k < NV; ++k) v[k] = out[k];
  __syncthreads();
}

// Block partials of this launch: part[k * gridDim.x + blockIdx.x].
template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double *part) {
  block_sum_all<NV>(v);
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; ++k) part[k * gridDim.x + blockIdx.x] = v[k];
}

// Fixed-order reduction of a producer's g block partials (same in every block).
template <int NV>
__device__ __forceinline__ void reduce_partials(const double *part, int g, double (&v)[NV]) {
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < g; b += kThreads)
    for (int k = 0; k < NV; ++k) v[k] += part[k * g + b];
  block_sum_all<NV>(v);
}

// logical block index giving each XCD (blocks dispatched round-robin over 8)
// a contiguous range
__device__ __forceinline__ int xcd_block() {
  const int g = gridDim.x, b = blockIdx.x;
  return (g % 8 == 0) ? (b % 8) * (g / 8) + b / 8 : b;
}

// Slice range of the calling wavefront, wave-uniform (scalar registers), so
// slice metadata and diagonal offsets are scalar loads.
// Two traversal orders (both give each XCD one contiguous eighth of the
// slices): order 0, each wave owns a contiguous run of slices; order 1, the
// XCD's waves sweep its eighth together (wave-strided), so the processing front
// is narrow and rows a few planes apart are touched close in time.
struct WaveRange {
  int s0, s1, step;
};
__device__ __forceinline__ WaveRange wave_slices(int ns, int order, int pstride = 0) {
  const int w = (int)(threadIdx.x >> 6);
  if (order == 2 && pstride > 0 && (int)gridDim.x * kWaves >= pstride) {   // every column has a wave
    // Pencils: wave (z, c) takes slices c + P j for j in its z-chunk, P = pstride
    // (about one plane), so the lower-part reads of the symmetric layout at
    // offset ~ -P slices touch lines the same wave loaded one step before.
    // Logical wave ids are contiguous per XCD, so an XCD owns whole z-chunks.
    const int P = pstride;
    const int wid = __builtin_amdgcn_readfirstlane(xcd_block() * kWaves + w);
    const int nw = gridDim.x * kWaves;
    const int np = (ns + P - 1) / P;                 // steps of P slices
    const int Z = max(1, nw / P);                    // z-chunks
    const int L = (np + Z - 1) / Z;                  // steps per chunk
    const int z = wid / P, c = wid % P;
    if (z >= Z) return {0, 0, 1};
    const int j0 = z * L, j1 = min(np, j0 + L);
    const int s0 = c + P * j0;
    const int s1 = min(ns, c + P * j1);
    return {s0, max(s0, s1), P};
  }
  if (order == 1 && gridDim.x % 8 == 0) {
    const int b = blockIdx.x, x = b % 8, per = gridDim.x / 8;
    const int wx = __builtin_amdgcn_readfirstlane((b / 8) * kWaves + w);
    const int c0 = (int)((int64_t)ns * x / 8), c1 = (int)((int64_t)ns * (x + 1) / 8);
    return {c0 + wx, c1, per * kWaves};
  }
  const int wave = __builtin_amdgcn_readfirstlane(xcd_block() * kWaves + w);
  const int nw = gridDim.x * kWaves;
  return {(int)((int64_t)ns * wave / nw), (int)((int64_t)ns * (wave + 1) / nw), 1};
}

// Streamed vector data of the FP64 vector kernels, loaded and stored
// non-temporally: the CG vectors then do not displace the matrix from the
// 256 MB Infinity Cache between iterations (the 128^3 FP64 symmetric copy is
// 230 MB; measured: SpMV inside the solve 59 us against 71 us, CG 21.2 ms
// against 24.3 ms).
template <typename T>
__device__ __forceinline__ T ld_nt(const T *p) {
  return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ void st_nt(T *p, T v) {
  __builtin_nontemporal_store(v, p);
}

// The next FP32 CG direction p' = beta p + z, rounded once (fused multiply-add).
__device__ __forceinline__ float dir_value(float beta, float p, float z) { return __builtin_fmaf(beta, p, z); }

// Right-hand vectors of the slice kernels.  operator()(c) is element c;
// off(o, rb) is element row + o where rb = row * sizeof(elem): a scalar base
// pointer per diagonal plus a per-lane 32-bit byte offset, so a gather needs no
// vector address arithmetic.
template <typename X>
struct VecRef {   // x
  using elem = X;
  const X *__restrict__ x;
  __device__ __forceinline__ X operator()(int c) const { return x[c]; }
  __device__ __forceinline__ X off(int o, uint32_t rb) const { return *(const X *)((const char *)(x + o) + rb); }
};
// FP64 solution plus FP32 increment (scaled by sc when set): the solution a
// reliable update makes.
__device__ __forceinline__ double new_x(double x, float d, const double *sc, int64_t i) {
  return sc ? __builtin_fma(sc[i], (double)d, x) : x + (double)d;
}

// Row `lane` of slice s (s wave-uniform): returns (A g)_row.  Slices whose
// columns may leave [0, n) (flag bit 29) clamp the column.
// Symmetric-diagonal rows: y_r = sum_k V[k][r] g[r+o_k] + sum_{k>0} V[k][r-o_k] g[r-o_k]
// (o_0 = 0).  Indices are clamped into [0, n) and out-of-range terms zeroed;
// padding values are 0.
//
// Column triplets (m.trip): the 14 stored offsets of a 27-point stencil on a
// structured grid are 0, 1 and the triplets {d-1, d, d+1} for d = mx, P-mx, P,
// P+mx (row length mx, plane P), so the 27 columns of a row are the triplets
// around row + d for d = 0, +-mx, +-(P-mx), +-P, +-(P+mx).  One gather per
// triplet loads its centre; the -1/+1 neighbours are the centres of the
// adjacent lanes (shuffles), and the two lanes at the slice's ends take theirs
// from wave-uniform loads.  That is 9 vector gathers of g instead of 27; the
// terms are summed in the generic order, so the result is the same bit for bit.
template <typename T, typename A, typename G>
__device__ __forceinline__ A sym27_row(const SlicedDev &m, const T *__restrict__ v, const G &g, int s, int lane) {
  using X = typename G::elem;
  const int row0 = s * kSlice;
  const int row = row0 + lane;
  const int hi = m.n - 1;
  X xv[9][3];
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int d = t == 0 ? 0 : (t <= 4 ? m.soffs[3 * t] : -m.soffs[3 * (t - 4)]);
    const X c = g(min(max(row + d, 0), hi));
    const X hl = g(__builtin_amdgcn_readfirstlane(min(max(row0 + d - 1, 0), hi)));
    const X hr = g(__builtin_amdgcn_readfirstlane(min(max(row0 + kSlice + d, 0), hi)));
    const X l = __shfl_up(c, 1, 64), r = __shfl_down(c, 1, 64);
    xv[t][0] = lane == 0 ? hl : l;
    xv[t][1] = c;
    xv[t][2] = lane == kSlice - 1 ? hr : r;
  }
  A acc = 0;
  // upper part: k = 0 (column row), 1 (row + 1), k >= 2: triplet (k+1)/3, member (k+1)%3
  if (m.unit) acc += (A)xv[0][1];
  else acc += (A)v[sym_at(0, row, m.stride, 14)] * (A)xv[0][1];
  acc += (A)v[sym_at(1, row, m.stride, 14)] * (A)xv[0][2];
#pragma unroll
  for (int k = 2; k < 14; ++k) acc += (A)v[sym_at(k, row, m.stride, 14)] * (A)xv[(k + 1) / 3][(k + 1) % 3];
  // lower part: column row - o_k: k = 1 -> row - 1; k >= 2 -> the mirrored triplet
#pragma unroll
  for (int k = 1; k < 14; ++k) {
    const int rr = row - m.soffs[k];
    const int rc = min(max(rr, 0), hi);
    A a = (A)v[sym_at(k, rc, m.stride, 14)];
    if (rr < 0) a = 0;
    const X xx = k == 1 ? xv[0][0] : xv[(k + 1) / 3 + 4][2 - (k + 1) % 3];
    acc += a * (A)xx;
  }
  return acc;
}

template <typename T, typename A, typename G>
__device__ __forceinline__ A sym_rows(const SlicedDev &m, const T *__restrict__ v, const G &g, int s, int lane) {
  if (m.trip) return sym27_row<T, A, G>(m, v, g, s, lane);
  const int row = s * kSlice + lane;
  const int hi = m.n - 1;
  A acc = 0;
  if (m.kplus == 14) {   // 27-point stencils: every load of the row in flight at once
#pragma unroll
    for (int k = 0; k < 14; ++k) {
      const int o = m.soffs[k];
      const A a = (k == 0 && m.unit) ? (A)1 : (A)v[sym_at(k, row, m.stride, 14)];
      const int c = min(row + o, hi);
      acc += a * (A)g(c);
    }
#pragma unroll
    for (int k = 1; k < 14; ++k) {
      const int o = m.soffs[k];
      const int rr = row - o;
      const int rc = min(max(rr, 0), hi);
      A a = (A)v[sym_at(k, rc, m.stride, 14)];
      if (rr < 0) a = 0;
      acc += a * (A)g(rc);
    }
    return acc;
  }
  for (int k = 0; k < m.kplus; ++k) {   // diagonal and upper part
    const int o = m.soffs[k];
    const A a = (k == 0 && m.unit) ? (A)1 : (A)v[sym_at(k, row, m.stride, m.kplus)];
    const int c = min(row + o, hi);
    acc += a * (A)g(c);
  }
  for (int k = 1; k < m.kplus; ++k) {   // lower part from the stored transpose
    const int o = m.soffs[k];
    const int rr = row - o;
    const int rc = min(max(rr, 0), hi);
    A a = (A)v[sym_at(k, rc, m.stride, m.kplus)];
    if (rr < 0) a = 0;
    acc += a * (A)g(rc);
  }
  return acc;
}

template <typename T, typename A, typename G, int U = 9>
__device__ __forceinline__ A slice_rows(const SlicedDev &m, const T *__restrict__ v, const G &g, int s, int lane) {
  if (m.sym) return sym_rows<T, A, G>(m, v, g, s, lane);
  const int64_t b = m.base[s];
  const int meta = m.meta[s];
  const int K = meta & 0x0fffffff;
  const int row = s * kSlice + lane;
  A acc = 0;
  const T *vv = v + b + lane;
  if (!(meta >> 30)) {
    const int32_t *off = m.offs + m.aux[s];
    if (!((meta >> 29) & 1)) {
      const uint32_t rb = (uint32_t)row * (uint32_t)sizeof(typename G::elem);
      if (K == 27) {   // 27-point stencil slices: all loads of the slice in flight at once
#pragma unroll
        for (int k = 0; k < 27; ++k) {
          const A a = (A)vv[k * kSlice];
          acc += a * (A)g.off(off[k], rb);
        }
        return acc;
      }
      // U loads of values and of x in flight: 25 in the CG SpMV (64^3 Q2 class
      // slices of 125, 75, 45 and 27 offsets: 386 us against 416 us with 9; 32
      // and more more than doubled the time); 9 in kernels held to 128 VGPRs
#pragma unroll U
      for (int k = 0; k < K; ++k) {
        const A a = (A)vv[k * kSlice];
        acc += a * (A)g.off(off[k], rb);
      }
    } else {
      const int hi = m.n - 1;
      for (int k = 0; k < K; ++k) {
        const A a = (A)vv[k * kSlice];
        const int c = min(max(row + off[k], 0), hi);
        acc += a * (A)g(c);
      }
    }
  } else {
    const int32_t *ec = m.ecol + m.aux[s] + lane;
#pragma unroll 4
    for (int k = 0; k < K; ++k) {
      const A a = (A)vv[k * kSlice];
      const int c = ec[k * kSlice];
      acc += a * (A)g(c);
    }
  }
  return acc;
}

template <typename T, typename A, int U = 9>
__device__ __forceinline__ A slice_row(const SlicedDev &m, const T *__restrict__ v, const T *__restrict__ x, int s,
                                       int lane) {
  return slice_rows<T, A, VecRef<T>, U>(m, v, VecRef<T>{x}, s, lane);
}

// ------------------------------------------------------------------ kernels ---

// r = b - A x (x_zero: r = b) ; z = M r ; p = z ; partials <r,z>, ||r||^2, ||b||^2, <b,Mb>
template <typename T>
__global__ __launch_bounds__(kThreads) void k_init(SlicedDev m, const T *__restrict__ v, const T *__restrict__ b,
                                                   const T *__restrict__ x, int x_zero, T *__restrict__ r,
                                                   T *__restrict__ p, const T *__restrict__ invd, double *part) {
  const int lane = threadIdx.x & 63;
  const WaveRange wr = wave_slices(m.ns, m.order, m.pstride);
  const int s0 = wr.s0, s1 = wr.s1, sstep = wr.step;
  double acc[4] = {0, 0, 0, 0};
  for (int s = s0; s < s1; s += sstep) {
    T ax = x_zero ? (T)0 : slice_row<T, T>(m, v, x, s, lane);
    const int row = slice_row_index(m, s, lane);
    if (row < m.n) {
      T bi = b[row], ri = bi - ax;
      T zi = invd ? invd[row] * ri : ri;
      r[row] = ri;
      p[row] = zi;
      T mb = invd ? invd[row] * bi : bi;
      acc[0] += (double)(ri * zi);
      acc[1] += (double)(ri * ri);
      acc[2] += (double)(bi * bi);
      acc[3] += (double)(bi * mb);
    }
  }
  write_partials<4>(acc, part);
}

// one block: init_convergence_data (cg.f90:279-321).  A raised zero-pivot flag
// of the Jacobi setup ends the solve before its first iteration.
__global__ __launch_bounds__(kThreads) void k_init_fin(const double *part, int g, CgState *S, CgResult *R,
                                                       const int *zero_diag) {
  double v[4];
  reduce_partials<4>(part, g, v);
  if (threadIdx.x != 0) return;
  CgState &s = S[0];
  const double rz = v[0], rnorm = sqrt(v[1]), bnorm = sqrt(v[2]), bnrm_m = sqrt(v[3]);
  const double rtol = s.rtol, atol = s.atol;
  const int c = s.crit;
  s.rz = rz;
  s.rnrm_m = sqrt(rz);
  s.bnorm = bnorm;
  s.r0norm = rnorm;
  s.rnorm = rnorm;
  double t = 0.0, extra = 0.0;
  switch (c) {
    case FEMHIP_DELTA_RHS:
    case FEMHIP_RES_RHS:
    case FEMHIP_DELTA_RHS_AND_RES_RHS:
    case FEMHIP_DELTA_DELTA_AND_RES_RHS:
      t = rtol * bnorm + atol;
      if (c == FEMHIP_DELTA_RHS_AND_RES_RHS || c == FEMHIP_DELTA_DELTA_AND_RES_RHS) extra = t;
      break;
    case FEMHIP_RES_RES:
    case FEMHIP_DELTA_DELTA_AND_RES_RES:
      t = rtol * rnorm + atol;
      if (c == FEMHIP_DELTA_DELTA_AND_RES_RES) extra = t;
      break;
    case FEMHIP_DELTA_RHS_AND_RES_RES:
      t = rtol * bnorm + atol;
      extra = rtol * rnorm + atol;
      break;
    case FEMHIP_RES_NRMGIVEN_RHS_NRMGIVEN: t = rtol * bnrm_m + atol; break;
    case FEMHIP_RES_NRMGIVEN_RES_NRMGIVEN: t = rtol * sqrt(rz) + atol; break;
    default: break;
  }
  s.rhs_test = t;
  s.rhs_extra = extra;
  s.iter = 0;
  s.done = (zero_diag && *zero_diag) ? 1 : 0;
  s.converged = 0;
  s.est = 0.0;
  s.ru = 0;
  s.updates = 0;
  s.rmax = rnorm;
  S[1] = s;
  R->est = 0.0;
  R->rhs_test = t;
  R->rnorm = rnorm;
  R->bnorm = bnorm;
  R->iter = 0;
  R->done = s.done;
  R->converged = 0;
  R->updates = 0;
  R->prog = 0.0;   // no estimate before the first iteration
  R->prog_it = 0;
}

// q = A p ; partials <p,q>, ||p||^2
template <typename T, typename A>
__global__ __launch_bounds__(kThreads) void k_spmv(SlicedDev m, const T *__restrict__ v, const T *__restrict__ p,
                                                   T *__restrict__ q, const CgState *S, int par, double *part,
                                                   int rev) {
  if (S[par].done) return;
  const int lane = threadIdx.x & 63;
  const WaveRange wr = wave_slices(m.ns, m.order, m.pstride);
  const int s0 = wr.s0, s1 = wr.s1, sstep = wr.step;
  A pq = 0, pp = 0;
  for (int t = s0; t < s1; t += sstep) {
    const int s = rev ? m.ns - 1 - t : t;
    const A acc = slice_row<T, A, 25>(m, v, p, s, lane);
    const int row = slice_row_index(m, s, lane);
    if (row < m.n) {
      q[row] = (T)acc;
      const A pi = (A)p[row];
      pq += pi * (A)(T)acc;
      pp += pi * pi;
    }
  }
  double red[2] = {(double)pq, (double)pp};
  write_partials<2>(red, part);
}

// alpha ; r -= alpha q ; partials <r,z>, ||r||^2 (x += alpha p is made by
// k_dir, which reads p anyway: 32 + 48 bytes a row for the two kernels
// instead of 56 + 32)
template <typename T, typename A>
__global__ __launch_bounds__(kThreads) void k_update(int32_t n, T *__restrict__ r, const T *__restrict__ q,
                                                     const T *__restrict__ invd, CgState *S, int par,
                                                     const double *part_s, int gs, double *part) {
  if (S[par].done) return;
  const int64_t i0 = (int64_t)blockIdx.x * kThreads + threadIdx.x, stride = (int64_t)gridDim.x * kThreads;
  T rv[kPrefetch], qv[kPrefetch], dv[kPrefetch];
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k) {   // in flight while the partials are reduced
    const int64_t i = i0 + k * stride;
    if (i < n) {
      rv[k] = ld_nt(&r[i]);
      qv[k] = ld_nt(&q[i]);
      dv[k] = invd ? ld_nt(&invd[i]) : (T)1;
    }
  }
  double red[2];
  reduce_partials<2>(part_s, gs, red);
  const double alpha = (red[0] != 0.0) ? S[par].rz / red[0] : 0.0;   // cg.f90:195-199
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S[par].alpha = alpha;
    S[par].pap = red[0];
    S[par].pp = red[1];
  }
  const T a = (T)alpha;
  A rz = 0, rr = 0;
  auto row = [&](int64_t i, T r0, T qi, T di) {
    const T ri = r0 + (-a) * qi;
    st_nt(&r[i], ri);
    const T zi = invd ? di * ri : ri;
    rz += (A)ri * (A)zi;
    rr += (A)ri * (A)ri;
  };
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k)
    if (i0 + k * stride < n) row(i0 + k * stride, rv[k], qv[k], dv[k]);
  for (int64_t i = i0 + kPrefetch * stride; i < n; i += stride) row(i, r[i], q[i], invd ? invd[i] : (T)1);
  double out[2] = {(double)rz, (double)rr};
  write_partials<2>(out, part);
}

// stopping test (cg.f90:323-407) ; beta ; x += alpha p ; p = z + beta p
// (the last iteration: x += alpha p only).  Rh (the last iteration of a chunk
// only): host-mapped copy of the result block, read by the host after the
// chunk, so no copy kernel is needed.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_dir(int32_t n, T *__restrict__ x, T *__restrict__ p,
                                                  const T *__restrict__ r, const T *__restrict__ invd, CgState *S,
                                                  int par, const double *part_u, int gu, double *hist, CgResult *R,
                                                  CgResult *Rh) {
  const CgState &cur = S[par];
  if (cur.done) {   // keep `done` sticky in both parity slots
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      S[par ^ 1].done = 1;
      if (Rh) *Rh = *R;
    }
    return;
  }
  const int64_t i0 = (int64_t)blockIdx.x * kThreads + threadIdx.x, stride = (int64_t)gridDim.x * kThreads;
  T pv[kPrefetch], xv[kPrefetch], rv[kPrefetch], dv[kPrefetch];
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k) {   // in flight while the partials are reduced
    const int64_t i = i0 + k * stride;
    if (i < n) {
      pv[k] = ld_nt(&p[i]);
      xv[k] = ld_nt(&x[i]);
      rv[k] = ld_nt(&r[i]);
      dv[k] = invd ? ld_nt(&invd[i]) : (T)1;
    }
  }
  double red[2];
  reduce_partials<2>(part_u, gu, red);
  const int it = cur.iter + 1;
  const int c = cur.crit;
  const double rnorm = sqrt(red[1]);
  double est = cur.est, rhs_test = cur.rhs_test;
  if (it == 1 && is_delta(c)) {
    est = cur.alpha * sqrt(cur.pp);
    if (is_delta_delta(c)) rhs_test = cur.rtol * est + cur.atol;
  }
  if (c == FEMHIP_RES_RES || c == FEMHIP_RES_RHS) est = rnorm;
  else if (is_delta(c)) {
    if (it != 1) est = cur.alpha * sqrt(cur.pp);
  } else est = cur.rnrm_m;
  int conv = est <= rhs_test;
  if (c >= FEMHIP_DELTA_RHS_AND_RES_RES && conv) conv = rnorm <= cur.rhs_extra;
  const bool done = conv || it >= cur.maxit;
  double beta = 1.0 / cur.rz;
  beta = beta * red[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CgState nx = cur;
    nx.iter = it;
    nx.est = est;
    nx.rhs_test = rhs_test;
    nx.rnorm = rnorm;
    nx.converged = conv;
    nx.done = done;
    nx.beta = beta;
    nx.rz = red[0];
    nx.rnrm_m = sqrt(red[0]);
    S[par ^ 1] = nx;
    if (hist) hist[it - 1] = est;
    R->iter = it;
    R->est = est;
    R->rhs_test = rhs_test;
    R->rnorm = rnorm;
    R->converged = conv;
    R->done = done;
    R->prog = est;
    R->prog_it = it;
    if (Rh) *Rh = *R;
  }
  const T a = (T)cur.alpha;
  if (done) {
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k)
      if (i0 + k * stride < n) x[i0 + k * stride] = xv[k] + a * pv[k];
    for (int64_t i = i0 + kPrefetch * stride; i < n; i += stride) x[i] = x[i] + a * p[i];
    return;
  }
  const T bt = (T)beta;
  auto row = [&](int64_t i, T pi, T xi, T ri, T di) {
    st_nt(&x[i], xi + a * pi);
    const T zi = invd ? di * ri : ri;
    const T t = bt * pi;
    st_nt(&p[i], t + zi);
  };
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k)
    if (i0 + k * stride < n) row(i0 + k * stride, pv[k], xv[k], rv[k], dv[k]);
  for (int64_t i = i0 + kPrefetch * stride; i < n; i += stride) row(i, p[i], x[i], r[i], invd ? invd[i] : (T)1);
}

// ------------------------------------------------- FP32 with reliable updates ---

// With the Jacobi preconditioner the FP32 iterations run on the symmetrically
// scaled system  (S A S) y = S r,  S = diag(A)^(-1/2)  (sc = S in FP64): CG on
// S A S is Jacobi-PCG on A in other variables (same Krylov spaces), its matrix
// has a unit diagonal that is not read, and z = r needs no preconditioner
// pass.  The increment d lives in the scaled variables; a reliable update
// folds x += S d into the FP64 solution and scales the FP64 residual b - A x
// into the FP32 one.  Without a preconditioner sc is null and S = I.

// r64 = b - A64 x64 ; r32 = S r64 ; p = r32 ; d = 0 ; partials <r32,r32>, ||r64||^2, ||b||^2, ||r32||^2
__global__ __launch_bounds__(kThreads) void k_ru_init(SlicedDev m, const double *__restrict__ v64,
                                                      const double *__restrict__ b, const double *__restrict__ x,
                                                      int x_zero, double *__restrict__ r64, float *__restrict__ r32,
                                                      float *__restrict__ p32, float *__restrict__ d32,
                                                      const double *__restrict__ sc, double *part) {
  const int lane = threadIdx.x & 63;
  const WaveRange wr = wave_slices(m.ns, m.order, m.pstride);
  const int s0 = wr.s0, s1 = wr.s1, sstep = wr.step;
  double acc[4] = {0, 0, 0, 0};
  for (int s = s0; s < s1; s += sstep) {
    const double ax = x_zero ? 0.0 : slice_row<double, double>(m, v64, x, s, lane);
    const int row = slice_row_index(m, s, lane);
    if (row < m.n) {
      const double bi = b[row], ri = bi - ax;
      r64[row] = ri;
      const float rf = (float)(sc ? sc[row] * ri : ri);
      r32[row] = rf;
      p32[row] = rf;
      d32[row] = 0.0f;
      acc[0] += (double)(rf * rf);
      acc[1] += ri * ri;
      acc[2] += bi * bi;
      acc[3] += (double)(rf * rf);
    }
  }
  write_partials<4>(acc, part);
}

__global__ __launch_bounds__(kThreads) void k_ru_init_fin(const double *part, int g, CgState *S, CgResult *R,
                                                          const int *zero_diag) {
  double v[4];
  reduce_partials<4>(part, g, v);
  if (threadIdx.x != 0) return;
  CgState &s = S[0];
  const double rnorm = sqrt(v[1]), bnorm = sqrt(v[2]);
  s.rz = v[0];
  s.bnorm = bnorm;
  s.r0norm = rnorm;
  s.rnorm = rnorm;
  s.rhs_test = (s.crit == FEMHIP_RES_RHS ? bnorm : rnorm) * s.rtol + s.atol;
  s.rmax = sqrt(v[3]);
  s.rscale = rnorm > 0.0 ? sqrt(v[3]) / rnorm : 1.0;
  s.iter = 0;
  s.converged = rnorm <= s.rhs_test;
  s.done = s.converged || (zero_diag && *zero_diag);
  s.ru = 0;
  s.updates = 0;
  s.first = 1;
  S[1] = s;
  R->est = rnorm;
  R->rhs_test = s.rhs_test;
  R->rnorm = rnorm;
  R->bnorm = bnorm;
  R->iter = 0;
  R->done = s.done;
  R->converged = s.converged;
  R->updates = 0;
  R->prog = rnorm;
  R->prog_it = 0;
}

// One iteration slot of the FP32 path is either a CG iteration or, when the
// previous slot requested it (state.ru), a reliable update; the same three
// kernels serve both, so an update costs no extra launches.

// slot 1: q = A32 p (partials <p,q>, ||p||^2), or r = b - A64 x in FP64 (the
// slot before folded the increment into x), scaled and rounded into r32
// (partials ||r64||^2, <r32,r32>, ||r32||^2).
// m32: the FP32 matrix's view (unit diagonal when scaled).  Two blocks per CU
// must stay resident for the FP32 iterations (4 waves per SIMD: at most 128
// VGPRs; the second bound is waves per execution unit); the rare
// FP64 update branch, which alone would take more, absorbs the limit.
// A: the type the FP32 matrix rows are summed in (double by default: 259
// inner iterations at 128^3 against 281 with float; float for FP32_STRICT).
template <typename A>
__global__ __launch_bounds__(kThreads, 4) void k_ru_spmv(SlicedDev m, SlicedDev m32, const float *__restrict__ v32,
                                                      const double *__restrict__ v64, const float *__restrict__ p,
                                                      float *__restrict__ q, const double *__restrict__ b,
                                                      const double *__restrict__ x, float *__restrict__ r32,
                                                      const double *__restrict__ sc, const CgState *S, int par,
                                                      double *part, int rev) {
  const CgState &cur = S[par];
  if (cur.done) return;
  const int lane = threadIdx.x & 63;
  const WaveRange wr = wave_slices(m.ns, m.order, m.pstride);
  const int s0 = wr.s0, s1 = wr.s1, sstep = wr.step;
  double acc[3] = {0, 0, 0};
  if (!cur.ru) {
    float pq = 0.f, pp = 0.f;
    for (int t = s0; t < s1; t += sstep) {
      const int s = rev ? m.ns - 1 - t : t;
      // FP32 values and vectors, the row sum held in FP64 registers: 8 % fewer
      // inner iterations at 128^3 (259 against 281), no extra memory traffic
      const float a = (float)slice_row<float, A>(m32, v32, p, s, lane);
      const int row = slice_row_index(m, s, lane);
      if (row < m.n) {
        q[row] = a;
        const float pi = p[row];
        pq += pi * a;
        pp += pi * pi;
      }
    }
    acc[0] = pq;
    acc[1] = pp;
  } else {
    for (int s = s0; s < s1; s += sstep) {
      const double ax = slice_row<double, double>(m, v64, x, s, lane);
      const int row = slice_row_index(m, s, lane);
      if (row < m.n) {
        const double ri = b[row] - ax;
        const float rf = (float)(sc ? sc[row] * ri : ri);
        r32[row] = rf;
        acc[0] += ri * ri;
        acc[1] += (double)(rf * rf);
        acc[2] += (double)(rf * rf);
      }
    }
  }
  write_partials<3>(acc, part);
}

// slot 2: alpha ; r -= alpha q (partials <r,r> twice: <r,z> and ||r||^2), or
// the update's second half: FP64 stopping test ; p = r + beta p
__global__ __launch_bounds__(kThreads) void k_ru_update(int32_t n, float *__restrict__ r, float *__restrict__ p,
                                                        const float *__restrict__ q, CgState *S, int par,
                                                        const double *part_s, int gs, double *part) {
  // Block 0 writes only fields the other blocks do not read (alpha, pap, pp,
  // ru_*); they read done, ru, rz, iter, maxit and rhs_test.
  const CgState &cur = S[par];
  if (cur.done) return;
  const int64_t i0 = (int64_t)blockIdx.x * kThreads + threadIdx.x, stride = (int64_t)gridDim.x * kThreads;
  float rv[kPrefetch], qv[kPrefetch];
  if (!cur.ru) {
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {   // in flight while the partials are reduced
      const int64_t i = i0 + k * stride;
      if (i < n) {
        rv[k] = r[i];
        qv[k] = q[i];
      }
    }
  }
  double red[3];
  reduce_partials<3>(part_s, gs, red);
  if (cur.ru) {
    const double rnorm = sqrt(red[0]);
    const int conv = rnorm <= cur.rhs_test;
    const int done = conv || cur.iter >= cur.maxit;
    double beta = 1.0 / cur.rz;
    beta = beta * red[1];
    const float bt = (float)beta;
    if (!done)
      for (int64_t i = i0; i < n; i += stride) p[i] = dir_value(bt, p[i], r[i]);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      CgState &w = S[par];
      w.ru_rz = red[1];
      w.ru_rmax = sqrt(red[2]);
      w.ru_rnorm = rnorm;
      w.ru_beta = beta;
      w.ru_conv = conv;
      w.ru_done = done;
      w.ru_rscale = rnorm > 0.0 ? sqrt(red[2]) / rnorm : cur.rscale;
    }
    return;
  }
  const double alpha = (red[0] != 0.0) ? cur.rz / red[0] : 0.0;   // cg.f90:195-199
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S[par].alpha = alpha;
    S[par].pap = red[0];
    S[par].pp = red[1];
  }
  const float a = (float)alpha;
  float rr = 0.f;
  auto row = [&](int64_t i, float r0, float qi) {
    const float ri = r0 + (-a) * qi;
    r[i] = ri;
    rr += ri * ri;
  };
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k)
    if (i0 + k * stride < n) row(i0 + k * stride, rv[k], qv[k]);
  for (int64_t i = i0 + kPrefetch * stride; i < n; i += stride) row(i, r[i], q[i]);
  double out[2] = {(double)rr, (double)rr};
  write_partials<2>(out, part);
}

// slot 3: d += alpha p, then take the CG direction step p = r + beta p, or
// request a reliable update and fold the increment into the FP64 solution
// (x += S d, d = 0); after an update, publish its results to the next parity
// slot
__global__ __launch_bounds__(kThreads) void k_ru_dir(int32_t n, float *__restrict__ d, float *__restrict__ p,
                                                     const float *__restrict__ r, double *__restrict__ x,
                                                     const double *__restrict__ sc, CgState *S, int par,
                                                     const double *part_u, int gu, CgResult *R, CgResult *Rh) {
  const CgState &cur = S[par];
  if (cur.done) {   // keep `done` sticky in both parity slots
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      S[par ^ 1].done = 1;
      if (Rh) *Rh = *R;
    }
    return;
  }
  if (cur.ru) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      CgState nx = cur;
      nx.ru = 0;
      nx.rz = cur.ru_rz;
      nx.rnrm_m = sqrt(cur.ru_rz);
      nx.rmax = cur.ru_rmax;
      nx.beta = cur.ru_beta;
      nx.est = cur.ru_rnorm;
      nx.rnorm = cur.ru_rnorm;
      nx.converged = cur.ru_conv;
      nx.done = cur.ru_done;
      nx.updates = cur.updates + 1;
      nx.rscale = cur.ru_rscale;
      S[par ^ 1] = nx;
      R->iter = cur.iter;
      R->est = cur.ru_rnorm;
      R->rnorm = cur.ru_rnorm;
      R->converged = cur.ru_conv;
      R->done = cur.ru_done;
      R->updates = nx.updates;
      R->prog = cur.ru_rnorm;
      R->prog_it = cur.iter;
      if (Rh) *Rh = *R;
    }
    return;
  }
  const int64_t i0 = (int64_t)blockIdx.x * kThreads + threadIdx.x, stride = (int64_t)gridDim.x * kThreads;
  float pv[kPrefetch], dv[kPrefetch], rv[kPrefetch];
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k) {   // in flight while the partials are reduced
    const int64_t i = i0 + k * stride;
    if (i < n) {
      pv[k] = p[i];
      dv[k] = d[i];
      rv[k] = r[i];
    }
  }
  double red[2];
  reduce_partials<2>(part_u, gu, red);
  const int it = cur.iter + 1;
  const double rnorm = sqrt(red[1]);   // of the scaled FP32 residual
  // residual fell by delta since the last update, looks converged (threshold
  // carried into the scaled norm by the last update's ratio), or budget spent
  const bool ru = rnorm < cur.delta * cur.rmax || rnorm <= cur.rhs_test * cur.rscale || it >= cur.maxit;
  double beta = 1.0 / cur.rz;
  beta = beta * red[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CgState nx = cur;
    nx.iter = it;
    nx.ru = ru;
    nx.rnorm = rnorm;
    if (!ru) {
      nx.beta = beta;
      nx.rz = red[0];
      nx.rmax = fmax(cur.rmax, rnorm);
    }
    S[par ^ 1] = nx;
    R->prog = cur.rscale > 0.0 ? rnorm / cur.rscale : rnorm;   // estimate in the stopping test's norm
    R->prog_it = it;
    if (Rh) *Rh = *R;
  }
  const float a = (float)cur.alpha;
  if (ru) {   // the next slot recomputes the residual of the new x
    auto fold = [&](int64_t i, float di, float pi) {
      x[i] = new_x(x[i], di + a * pi, sc, i);
      d[i] = 0.0f;
    };
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k)
      if (i0 + k * stride < n) fold(i0 + k * stride, dv[k], pv[k]);
    for (int64_t i = i0 + kPrefetch * stride; i < n; i += stride) fold(i, d[i], p[i]);
    return;
  }
  const float bt = (float)beta;
  auto row = [&](int64_t i, float pi, float di, float ri) {
    d[i] = di + a * pi;
    p[i] = dir_value(bt, pi, ri);
  };
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k)
    if (i0 + k * stride < n) row(i0 + k * stride, pv[k], dv[k], rv[k]);
  for (int64_t i = i0 + kPrefetch * stride; i < n; i += stride) row(i, p[i], d[i], r[i]);
}

// ---- two kernels per iteration slot ----
// The slot's direction step and product come from a recurrence: with
// p' = r + beta p, the product q' = A p' = A r + beta q, so one kernel can take
// the stopping/update decision, move x, p and d and form q' from an SpMV of r
// (final after the previous slot), and a second one makes the r update:
//   slot A: decision (reduces slot B's partials) ; d += alpha p ;
//           p = r + beta p ; q = A r + beta q ; partials <p,q>, ||p||^2
//           or, when a reliable update is due, x += S (d + alpha p) ; d = 0
//   slot B: alpha = rz/<p,q> ; r -= alpha q ; partials <r,r>
//           or, after a fold, r = S (b - A64 x) in FP64 ; partials <r,r>, ||b - A x||^2
// The bytes per row are those of the three-kernel form (40 in FP32); the
// third launch and its fixed cost go.  Both kernels run on the matrix grid
// over the same slices, so rows are visited as in the SpMV.
// State parity: slot A reads S[par] and writes S[par ^ 1]; slot B reads
// S[par ^ 1], its block 0 stores alpha there for the next slot A.

constexpr int kRfbGroup = 8;   // slices of a wave whose r and q loads k_rf_b keeps in flight together

template <typename A>
__global__ __launch_bounds__(kThreads, 4) void k_rf_a(SlicedDev m, SlicedDev m32, const float *__restrict__ v32,
                                                   const float *__restrict__ r, float *__restrict__ p,
                                                   float *__restrict__ q, float *__restrict__ d,
                                                   double *__restrict__ x, const double *__restrict__ sc,
                                                   CgState *S, int par, const double *part_b, int gb, double *part,
                                                   CgResult *R, CgResult *Rh, int rev) {
  const CgState &cur = S[par];
  if (cur.done) {   // keep `done` sticky in both parity slots
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      S[par ^ 1].done = 1;
      if (Rh) *Rh = *R;
    }
    return;
  }
  double red[2] = {0.0, 0.0};
  if (!cur.first) reduce_partials<2>(part_b, gb, red);
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (lead) {
    S[par ^ 1] = cur;
    S[par ^ 1].first = 0;
  }
  int fold = 0, upd_d = 0, done = 0;
  double beta = 0.0;
  if (cur.first) {           // p = r, q = A r
  } else if (!cur.ru) {      // an iteration's r update has been made
    const int it = cur.iter + 1;
    const double rnorm = sqrt(red[0]);   // of the scaled FP32 residual
    // residual fell by delta since the last update, looks converged (threshold
    // carried into the scaled norm by the last update's ratio), or budget spent
    fold = rnorm < cur.delta * cur.rmax || rnorm <= cur.rhs_test * cur.rscale || it >= cur.maxit;
    if (!fold) {
      beta = 1.0 / cur.rz;
      beta = beta * red[0];
      upd_d = 1;
    }
    if (lead) {
      CgState &w = S[par ^ 1];
      w.iter = it;
      w.rnorm = rnorm;
      w.ru = fold;
      if (!fold) {
        w.beta = beta;
        w.rz = red[0];
        w.rmax = fmax(cur.rmax, rnorm);
      }
      R->iter = it;
      R->prog = cur.rscale > 0.0 ? rnorm / cur.rscale : rnorm;   // estimate in the stopping test's norm
      R->prog_it = it;
    }
  } else {                   // slot B replaced r by the FP64 residual of the folded x
    const double rn64 = sqrt(red[1]);
    const int conv = rn64 <= cur.rhs_test;
    beta = 1.0 / cur.rz;
    beta = beta * red[0];
    done = conv || cur.iter >= cur.maxit;
    if (lead) {
      CgState &w = S[par ^ 1];
      w.ru = 0;
      w.rz = red[0];
      w.beta = beta;
      w.rmax = sqrt(red[0]);
      w.rscale = rn64 > 0.0 ? sqrt(red[0]) / rn64 : cur.rscale;
      w.updates = cur.updates + 1;
      w.est = rn64;
      w.rnorm = rn64;
      w.converged = conv;
      w.done = done;
      R->iter = cur.iter;
      R->est = rn64;
      R->rnorm = rn64;
      R->converged = conv;
      R->done = done;
      R->updates = cur.updates + 1;
      R->prog = rn64;
      R->prog_it = cur.iter;
    }
  }
  if (lead && Rh) *Rh = *R;
  if (done) return;
  const int lane = threadIdx.x & 63;
  const WaveRange wr = wave_slices(m.ns, m.order, m.pstride);
  const int s0 = wr.s0, s1 = wr.s1, sstep = wr.step;
  const float a = (float)cur.alpha;
  if (fold) {   // the next slot B recomputes the residual of the new x
    for (int s = s0; s < s1; s += sstep) {
      const int row = slice_row_index(m, s, lane);
      if (row < m.n) {
        x[row] = new_x(x[row], d[row] + a * p[row], sc, row);
        d[row] = 0.0f;
      }
    }
    return;
  }
  const float bt = (float)beta;
  const bool first = cur.first;
  float pq = 0.f, pp = 0.f;
  for (int t = s0; t < s1; t += sstep) {
    const int s = rev ? m.ns - 1 - t : t;
    // FP32 values and vectors, the row sum held in FP64 registers
    const float w = (float)slice_row<float, A>(m32, v32, r, s, lane);
    const int row = slice_row_index(m, s, lane);
    if (row < m.n) {
      const float ri = r[row];
      float pi = ri, qi = w;
      if (!first) {
        const float po = p[row];
        if (upd_d) d[row] = d[row] + a * po;
        pi = dir_value(bt, po, ri);
        qi = __builtin_fmaf(bt, q[row], w);
      }
      p[row] = pi;
      q[row] = qi;
      pq += pi * qi;
      pp += pi * pi;
    }
  }
  double out[2] = {(double)pq, (double)pp};
  write_partials<2>(out, part);
}

// SYM: the symmetric-diagonal copy, whose slice s holds rows 64 s .. 64 s + 63
// (no row index is loaded, so nothing but r and q is waited for)
template <bool SYM>
__global__ __launch_bounds__(kThreads, 4) void k_rf_b(SlicedDev m, const double *__restrict__ v64,
                                                   const double *__restrict__ b, const double *__restrict__ x,
                                                   float *__restrict__ r, const float *__restrict__ q,
                                                   const double *__restrict__ sc, CgState *S, int par,
                                                   const double *part_a, int ga, double *part) {
  CgState &st = S[par ^ 1];
  if (st.done) return;
  const int lane = threadIdx.x & 63;
  const WaveRange wr = wave_slices(m.ns, m.order, m.pstride);
  const int s0 = wr.s0, s1 = wr.s1, sstep = wr.step;
  double acc[2] = {0, 0};
  if (st.ru) {
    for (int s = s0; s < s1; s += sstep) {
      const double ax = slice_row<double, double>(m, v64, x, s, lane);
      const int row = slice_row_index(m, s, lane);
      if (row < m.n) {
        const double ri = b[row] - ax;
        const float rf = (float)(sc ? sc[row] * ri : ri);
        r[row] = rf;
        acc[0] += (double)(rf * rf);
        acc[1] += ri * ri;
      }
    }
    write_partials<2>(acc, part);
    return;
  }
  // the wave's slices kRfbGroup at a time, all their loads of r and q issued
  // before the first is used; the first group's are in flight while the
  // partials are reduced
  int rows[kRfbGroup];
  float rv[kRfbGroup], qv[kRfbGroup];
  auto load = [&](int sb) {
#pragma unroll
    for (int k = 0; k < kRfbGroup; ++k) {   // row indices first: on the general layouts they are loads
      const int s = sb + k * sstep;
      rows[k] = s >= s1 ? m.n : SYM ? s * kSlice + lane : slice_row_index(m, s, lane);
    }
#pragma unroll
    for (int k = 0; k < kRfbGroup; ++k) {
      rv[k] = 0.0f;
      qv[k] = 0.0f;
      if (rows[k] < m.n) {
        rv[k] = r[rows[k]];
        qv[k] = q[rows[k]];
      }
    }
  };
  load(s0);
  double red[2];
  reduce_partials<2>(part_a, ga, red);
  const double alpha = (red[0] != 0.0) ? st.rz / red[0] : 0.0;   // cg.f90:195-199
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st.alpha = alpha;
    st.pap = red[0];
    st.pp = red[1];
  }
  const float a = (float)alpha;
  float rr = 0.f;
  for (int sb = s0; sb < s1; sb += kRfbGroup * sstep) {
    if (sb != s0) load(sb);
    // every new value before the first store, so the stores wait for nothing
#pragma unroll
    for (int k = 0; k < kRfbGroup; ++k) rv[k] = rv[k] + (-a) * qv[k];
#pragma unroll
    for (int k = 0; k < kRfbGroup; ++k)
      if (rows[k] < m.n) {
        r[rows[k]] = rv[k];
        rr += rv[k] * rv[k];
      }
  }
  acc[0] = rr;
  write_partials<2>(acc, part);
}

// ------------------------------------------------------------------ helpers ---

// Jacobi setup: invD = 1 / diag(A) (csr_sparse_matrix.f90:4231-4256,
// serial_scalar_array.f90:581-588), or with `root` the scaling S = diag(A)^(-1/2)
// of the FP32 iterations, which needs a positive diagonal.  A zero pivot (or a
// negative one under `root`) raises the flag: 1 zero, 2 negative.
__device__ __forceinline__ void jacobi_store(double d, int root, double *out, int32_t i, int *flag) {
  if (d == 0.0) *flag = 1;
  else if (root && d < 0.0) *flag = 2;
  out[i] = root ? 1.0 / sqrt(d) : 1.0 / d;
}
__global__ void k_jacobi(int32_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                         const double *__restrict__ val, double *__restrict__ out, int root, int *flag) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t lo = rowptr[i], hi = rowptr[i + 1] - 1;
  double d = 0.0;
  while (lo <= hi) {
    int32_t mid = (lo + hi) >> 1;
    int32_t c = col[mid];
    if (c == i) {
      d = val[mid];
      break;
    }
    if (c < i) lo = mid + 1;
    else hi = mid - 1;
  }
  jacobi_store(d, root, out, i, flag);
}

// The same from the sliced copy, for operators held without a CSR: the entry
// of offset 0 of a diagonal slice, or the first ELL entry whose column is the
// row (real entries precede the padding).
// One thread per (slice, lane); a lane's row comes from the slice's row list
// when the slice holds a length-sorted window.
__global__ void k_jacobi_sliced(SlicedDev m, const double *__restrict__ v, double *__restrict__ out, int root,
                                int *flag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)m.ns * kSlice) return;
  const int32_t s = (int32_t)(t / kSlice), lane = (int32_t)(t % kSlice);
  const int32_t i = slice_row_index(m, s, lane);
  if (i >= m.n) return;
  const int32_t meta = m.meta[s];
  const int32_t K = meta & ((1 << 28) - 1);
  const bool ell = (meta >> 30) & 1;
  const int64_t b = m.base[s];
  const int32_t ax = m.aux[s];
  double d = 0.0;
  for (int32_t k = 0; k < K; ++k) {
    const bool hit = ell ? m.ecol[ax + (int64_t)k * kSlice + lane] == i : m.offs[ax + k] == 0;
    if (hit) {
      d = v[b + (int64_t)k * kSlice + lane];
      break;
    }
  }
  jacobi_store(d, root, out, i, flag);
}

// Symmetric-diagonal layout from the CSR: the entries of row r with col >= r
// go to V[k][r] for col - r = soffs[k] (both ascending); missing diagonals and
// rows in [n, stride) get 0.  One 64-thread block per 64 rows: the rows' CSR
// run (at most 64 x maxlen entries, maxlen = 2 kplus - 1 distinct offsets) is
// read coalesced into LDS, then each thread walks its row there and every
// store of one diagonal is coalesced.
constexpr int kFillRows = 64;
__global__ __launch_bounds__(kFillRows) void k_fill_sym(int32_t n, const int32_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ col,
                                                        const double *__restrict__ val,
                                                        const int32_t *__restrict__ soffs, int kplus, int64_t stride,
                                                        double *__restrict__ v64, float *__restrict__ v32) {
  extern __shared__ double fill_lds[];
  const int maxlen = 2 * kplus - 1;
  double *sv = fill_lds;
  int32_t *sc = (int32_t *)(fill_lds + kFillRows * maxlen);
  const int64_t r0 = (int64_t)blockIdx.x * kFillRows;
  const int lane = threadIdx.x;
  const int64_t r = r0 + lane;
  int32_t e0 = 0, cnt = 0;
  if (r0 < n) {
    e0 = rowptr[r0];
    cnt = rowptr[min(r0 + kFillRows, (int64_t)n)] - e0;
  }
  for (int j = lane; j < cnt; j += kFillRows) {
    sc[j] = col[e0 + j];
    sv[j] = val[e0 + j];
  }
  __syncthreads();
  int32_t j = 0, e = 0;
  if (r < n) {
    j = rowptr[r] - e0;
    e = rowptr[r + 1] - e0;
    while (j < e && sc[j] < r) ++j;
  }
  for (int k = 0; k < kplus; ++k) {
    double a = 0.0;
    if (j < e && sc[j] - (int32_t)r == soffs[k]) a = sv[j++];
    if (v64) v64[sym_at(k, r, stride, kplus)] = a;
    if (v32) v32[sym_at(k, r, stride, kplus)] = (float)a;
  }
}

// Jacobi setup from the symmetric-diagonal copy: diagonal 0 is offset 0.
// A zero pivot raises the flag (mapped host memory, read after the solve).
__global__ void k_jacobi_sym(int32_t n, const double *__restrict__ v, int64_t stride, int kplus,
                             double *__restrict__ out, int root, int *flag) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  jacobi_store(v[sym_at(0, i, stride, kplus)], root, out, i, flag);
}

// FP32 copy of the solver matrix, of S A S when sc (= S) is set:
// v32 = v64 * (sc[row] * sc[col]) (the product in that order keeps the copy
// symmetric); out-of-range columns carry 0 and are clamped.  One thread per
// (slice, lane).
__global__ void k_scale32(SlicedDev m, const double *__restrict__ v64, const double *__restrict__ sc,
                          float *__restrict__ v32) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int s = (int)(t / kSlice), lane = (int)(t % kSlice);
  if (s >= m.ns) return;
  const int hi = m.n - 1;
  if (m.sym) {
    const int row = s * kSlice + lane;
    for (int k = 0; k < m.kplus; ++k) {
      const int64_t e = sym_at(k, row, m.stride, m.kplus);
      double a = row < m.n ? v64[e] : 0.0;
      if (sc && row < m.n) a *= sc[row] * sc[min(row + m.soffs[k], hi)];
      v32[e] = (float)a;
    }
    return;
  }
  const int64_t b = m.base[s];
  const int meta = m.meta[s];
  const int K = meta & 0x0fffffff;
  const bool ell = (meta >> 30) & 1;
  const int row = slice_row_index(m, s, lane);
  for (int k = 0; k < K; ++k) {
    const int64_t e = b + (int64_t)k * kSlice + lane;
    double a = v64[e];
    if (sc && row < m.n) {
      const int c = ell ? m.ecol[m.aux[s] + (int64_t)k * kSlice + lane] : row + m.offs[m.aux[s] + k];
      a *= sc[row] * sc[min(max(c, 0), hi)];
    }
    v32[e] = (float)a;
  }
}

// operator order <-> solver order of a permuted layout: to[perm[i]] = from[i]
// (padding rows of `to` cleared first), or to[i] = from[perm[i]]
__global__ void k_permute_in(int32_t n, const int32_t *__restrict__ perm, const double *__restrict__ from,
                             double *__restrict__ to) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    to[perm[i]] = from[i];
}
__global__ void k_permute_out(int32_t n, const int32_t *__restrict__ perm, const double *__restrict__ from,
                              double *__restrict__ to) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    to[i] = from[perm[i]];
}

// gather CSR values into the sliced layout (FP64 and/or FP32)
__global__ void k_fill_sliced(int64_t E, const int32_t *__restrict__ map, const double *__restrict__ val,
                              double *__restrict__ v64, float *__restrict__ v32) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t k = map[e];
    const double a = k >= 0 ? val[k] : (k == -2 ? 1.0 : 0.0);   // -2: identity row of a padded layout
    if (v64) v64[e] = a;
    if (v32) v32[e] = (float)a;
  }
}

__global__ void k_fill0(int64_t n, double *__restrict__ a) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.0;
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// partials of ||b - A x||^2 (reporting only)
__global__ __launch_bounds__(kThreads) void k_true_res(SlicedDev m, const double *__restrict__ v64,
                                                       const double *__restrict__ b, const double *__restrict__ x,
                                                       double *part) {
  const int lane = threadIdx.x & 63;
  const WaveRange wr = wave_slices(m.ns, m.order, m.pstride);
  const int s0 = wr.s0, s1 = wr.s1, sstep = wr.step;
  double acc[1] = {0};
  for (int s = s0; s < s1; s += sstep) {
    const double ax = slice_row<double, double>(m, v64, x, s, lane);
    const int row = slice_row_index(m, s, lane);
    if (row < m.n) {
      const double ri = b[row] - ax;
      acc[0] += ri * ri;
    }
  }
  write_partials<1>(acc, part);
}

}  // namespace
}  // namespace femhip

using namespace femhip;

struct femhip_solver {
  femhip_ctx *ctx = nullptr;
  femhip_operator *op = nullptr;
  femhip_cg_params_t prm{};
  int32_t n = 0;
  int g_mat = 8, g_vec = 8;      // grid sizes of the matrix and vector kernels
  DevBuf<double> x, b, r, p, q, invd, part_i, part_s, part_u, part_r, hist;
  DevBuf<float> r32, p32, q32, d32;
  DevBuf<CgState> st;            // S[2]
  DevBuf<CgResult> res;
  int *h_flag = nullptr, *d_flag = nullptr;    // zero-pivot flag of the Jacobi setup (pinned, mapped)
  CgResult *h_res[2] = {nullptr, nullptr};     // pinned, device-mapped result slots
  CgResult *d_res_map[2] = {nullptr, nullptr}; // their device addresses
  hipEvent_t ev[2] = {nullptr, nullptr};       // completion of the chunk writing each slot
  hipGraphExec_t graphs[2][2] = {};            // [chunk size][result slot]
  int graph_kind = -1;
  bool stream_launches = false;  // never build chunk graphs (set by the request server)
  bool borrowed_host = false;    // h_res, h_flag and ev belong to the caller (solver_use_host_slots)
  bool prepared = false;
  bool have_solution = false;
  ~femhip_solver() {
    if (!borrowed_host) {
      for (auto *h : h_res)
        if (h) (void)hipHostFree(h);
      if (h_flag) (void)hipHostFree(h_flag);
      for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    }
    for (auto &row : graphs)
      for (auto g : row)
        if (g) (void)hipGraphExecDestroy(g);
  }
};

namespace femhip {
namespace {

SlicedDev sliced_dev(femhip_operator &op) {
  SlicedDev m;
  m.ns = op.sl.ns;
  m.n = op.sl.rows > 0 ? op.sl.rows : op.n;   // solver rows
  m.base = op.sl.base.p;
  m.meta = op.sl.meta.p;
  m.aux = op.sl.aux.p;
  m.offs = op.sl.offs.p;
  m.ecol = op.sl.ecol.p;
  m.sym = op.sl.sym ? 1 : 0;
  m.kplus = op.sl.kplus;
  m.stride = op.sl.stride;
  m.soffs = op.sl.soffs.p;
  m.srow = op.sl.srow.n > 1 ? op.sl.srow.p : nullptr;   // null: no permuted slice, rows need no load
  // the CG SpMVs of odd iterations sweep the slices in reverse, so each starts
  // on the part of the matrix the previous one read last, which is the part
  // still in the Infinity Cache (FEMHIP_SPMV_ALT=0: every sweep forward).
  // FP32-refined CG: 141.5 -> 134.8 ms at 192^3 (random field), 85.4 -> 81.9 ms
  // at 64^3 Q2
  static const char *alt = getenv("FEMHIP_SPMV_ALT");
  m.alt = alt ? atoi(alt) : 1;
  // symmetric layout: pencils of one plane (the largest stored offset) per step
  // for order 2; measured inside CG solves order 2 and order 1 run the SpMV
  // equally fast and order 1 gives the faster FP32 iterations, so 1 is default
  m.pstride = op.sl.sym && !op.sl.h_soffs.empty() ? std::max(1, op.sl.h_soffs.back() / kSlice) : 0;
  static const char *ord = getenv("FEMHIP_SLICE_ORDER");
  m.order = ord ? atoi(ord) : 1;
  // offsets 0, 1, then four triplets {d-1, d, d+1} (FEMHIP_TRIPLETS=0 disables)
  static const char *tri = getenv("FEMHIP_TRIPLETS");
  const std::vector<int32_t> &o = op.sl.h_soffs;
  bool trip = op.sl.sym && op.sl.kplus == 14 && (int)o.size() == 14 && o[0] == 0 && o[1] == 1 &&
              !(tri && atoi(tri) == 0);
  for (int t = 1; trip && t <= 4; ++t) trip = o[3 * t - 1] + 1 == o[3 * t] && o[3 * t] + 1 == o[3 * t + 1];
  m.trip = trip ? 1 : 0;
  m.unit = 0;
  return m;
}

bool mixed(const femhip_solver &s) { return s.prm.precision != FEMHIP_PRECISION_FP64; }
// FP32_STRICT: FP32 matrix integrated in FP32 on the matrix cores where that
// applies, FP32 row sums in the inner iterations
bool strict(const femhip_solver &s) { return s.prm.precision == FEMHIP_PRECISION_FP32_STRICT; }
// FP32 row sums in the refined iterations: always for FP32_STRICT, for
// FP32_REFINED when FEMHIP_RF_SUM32=1 (faster slots, more inner iterations at
// 128^3 Q1, fewer milliseconds on large high-contrast fields).  Read once.
bool sum32(const femhip_solver &s) {
  static const bool env = [] {
    const char *e = getenv("FEMHIP_RF_SUM32");
    return e && atoi(e) == 1;
  }();
  return strict(s) || env;
}

// The FP32 matrix as the kernels see it: with Jacobi it holds S A S, whose unit
// diagonal the symmetric layout does not read.
SlicedDev fp32_view(const femhip_solver &s, SlicedDev m) {
  m.unit = (s.prm.preconditioner == FEMHIP_PREC_JACOBI && m.sym) ? 1 : 0;
  return m;
}

// The refined solver's FP32 matrix is the FP64 one rounded (k_scale32) unless
// FEMHIP_FP32_INTEGRATION=1 asks for it integrated in FP32 on the matrix cores
// (assemble_sym32, scalar Q1).  Rounding is the default because it was
// measured faster: the integrated matrix carries a few FP32 ulps of
// accumulation error against the half ulp of rounding, and the reliable
// updates, which compare against the FP64 residual, then take more inner
// iterations (128^3 Q1: 336 against 263; 192^3 random field: 940 against
// 842).  Read per call.
bool fp32_integration() {
  const char *e = getenv("FEMHIP_FP32_INTEGRATION");
  return e && atoi(e) == 1;
}

void prepare(femhip_solver &s) {
  femhip_operator &op = *s.op;
  FH_CHECK(op.computed, FEMHIP_ERR_STATE, "solve before fe_affine_operator%compute");
  hipStream_t st = s.ctx->stream;
  const bool jac = s.prm.preconditioner == FEMHIP_PREC_JACOBI;
  const bool mx = mixed(s);
  if (!op.sl.built) build_sliced(op);
  const int32_t n = op.sl.rows > 0 ? op.sl.rows : op.n;   // solver rows (padded when permuted)
  if (op.sl.sym && op.gather == 1 && op.sym_gather == 0) setup_sym_gather(op);
  if (!s.prepared) {
    s.n = n;
    const int ns = op.sl.ns;
    // Blocks of 512 per CU, grids a multiple of the 8 XCDs.  Matrix kernels:
    // each XCD's waves sweep its eighth of the slices together, and the
    // symmetric layout re-reads a stored diagonal one plane of rows later; the
    // sweep is fastest when its front (resident waves x one slice of values)
    // is about 2 MB, well inside the XCD's 4 MB L2: measured at 128^3, 1 block
    // per CU for the FP64 values (SpMV 54 us against 67 us at 4) and 2 for the
    // FP32 values; the sliced formats (no re-read) take 2 for both (Q2 64^3:
    // FP64 SpMV 370 us at 2, 374 at 1, 380 at 3, 385 at 4, 414 at 8).
    // Vector kernels: 4 blocks per CU for FP64, 2 for FP32.  FEMHIP_CG_BPC
    // sets both, FEMHIP_CG_BPC_MAT the matrix kernels alone.
    const char *bpc_env = getenv("FEMHIP_CG_BPC");
    const char *mat_env = getenv("FEMHIP_CG_BPC_MAT");
    const int bpc = bpc_env ? std::max(1, atoi(bpc_env)) : (mx ? 2 : 4);
    const int bpc_mat = mat_env ? std::max(1, atoi(mat_env)) : bpc_env ? bpc : (mx || !op.sl.sym ? 2 : 1);
    const int cap = std::max(8, bpc * std::max(1, s.ctx->num_cus));
    const int cap_mat = std::max(8, bpc_mat * std::max(1, s.ctx->num_cus));
    s.g_mat = std::max(8, std::min(cap_mat, (ns + 2 * kWaves - 1) / (2 * kWaves)));
    s.g_mat = (s.g_mat + 7) / 8 * 8;
    s.g_vec = std::max(8, std::min(cap, (n + 2 * kThreads - 1) / (2 * kThreads)));
    s.x.alloc(n); s.b.alloc(n); s.r.alloc(n); s.p.alloc(n); s.q.alloc(n);
    s.part_i.alloc((size_t)4 * s.g_mat);
    s.part_s.alloc((size_t)3 * s.g_mat);
    s.part_u.alloc((size_t)2 * std::max(s.g_vec, s.g_mat));   // also the two-kernel slot B's partials
    s.part_r.alloc((size_t)3 * s.g_mat);
    s.hist.alloc((size_t)std::max(1, s.prm.max_num_iterations));
    s.st.alloc(2);
    s.res.alloc(1);
    if (jac) s.invd.alloc(n);   // invD (FP64) or S (FP32_REFINED)
    if (mx) {
      s.r32.alloc(n); s.p32.alloc(n); s.q32.alloc(n); s.d32.alloc(n);
    }
    for (int k = 0; k < 2; ++k) {   // written by the GPU, read by the host: mapped, coherent
      if (!s.borrowed_host)
        FH_HIP(hipHostMalloc((void **)&s.h_res[k], sizeof(CgResult), hipHostMallocMapped | hipHostMallocCoherent));
      std::memset(s.h_res[k], 0, sizeof(CgResult));
      FH_HIP(hipHostGetDevicePointer((void **)&s.d_res_map[k], s.h_res[k], 0));
    }
    if (!s.borrowed_host)
      FH_HIP(hipHostMalloc((void **)&s.h_flag, sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
    *s.h_flag = 0;
    FH_HIP(hipHostGetDevicePointer((void **)&s.d_flag, s.h_flag, 0));
    if (!s.borrowed_host)
      for (auto &e : s.ev) FH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    s.prepared = true;
  }
  // matrix-dependent work, redone whenever the operator was recomputed
  if (op.sl.filled_version != op.version) {
    if (!op.sl.v64.p) op.sl.v64.alloc((size_t)std::max<int64_t>(1, op.sl.entries));
    if (op.sl.sym) {   // stride is a multiple of kFillRows; every row has at most 2 kplus - 1 entries
      const size_t lds = (size_t)kFillRows * (2 * op.sl.kplus - 1) * (sizeof(double) + sizeof(int32_t));
      hipLaunchKernelGGL(k_fill_sym, dim3((unsigned)(op.sl.stride / kFillRows)), dim3(kFillRows), lds, st, op.n,
                         op.rowptr.p, op.col.p, op.val.p, op.sl.soffs.p, op.sl.kplus, op.sl.stride, op.sl.v64.p,
                         nullptr);
    } else {
      hipLaunchKernelGGL(k_fill_sliced, dim3(2048), dim3(256), 0, st, op.sl.entries, op.sl.map.p, op.val.p,
                         op.sl.v64.p, nullptr);
    }
    FH_HIP(hipGetLastError());
    op.sl.filled_version = op.version;
  }
  // Jacobi setup without a host round trip: invD for FP64, the scaling
  // S = diag^(-1/2) for the FP32 iterations; a bad pivot raises s.h_flag,
  // which ends the solve at its initial state and is reported by read_result
  if (jac && n > 0) {
    *s.h_flag = 0;
    const dim3 G((n + 255) / 256), B(256);
    const int root = mx ? 1 : 0;
    if (op.sl.sym)
      hipLaunchKernelGGL(k_jacobi_sym, G, B, 0, st, n, op.sl.v64.p, op.sl.stride, op.sl.kplus, s.invd.p, root,
                         s.d_flag);
    else if (op.structured || op.sl.permuted)   // no CSR in the solver's order: the diagonal of the sliced copy
      hipLaunchKernelGGL(k_jacobi_sliced, dim3((unsigned)(((int64_t)op.sl.ns * kSlice + 255) / 256)), B, 0, st,
                         sliced_dev(op), op.sl.v64.p, s.invd.p, root, s.d_flag);
    else
      hipLaunchKernelGGL(k_jacobi, G, B, 0, st, n, op.rowptr.p, op.col.p, op.val.p, s.invd.p, root, s.d_flag);
    FH_HIP(hipGetLastError());
  }
  // FP32 copy: of S A S with Jacobi, of A without
  const bool int32 = strict(s) || fp32_integration();
  if (mx && (op.sl.filled_version32 != op.version || op.sl.scaled32 != jac || op.sl.wants_int32 != int32)) {
    if (!op.sl.v32.p) op.sl.v32.alloc((size_t)std::max<int64_t>(1, op.sl.entries));
    // rounded from the FP64 copy, or integrated in FP32 on the matrix cores
    // where asked for and supported (scalar Q1 on the symmetric-diagonal copy)
    if (!int32 || !assemble_sym32(op, jac ? s.invd.p : nullptr, op.sl.v32.p)) {
      const int64_t threads = (int64_t)op.sl.ns * kSlice;
      hipLaunchKernelGGL(k_scale32, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, sliced_dev(op),
                         op.sl.v64.p, jac ? s.invd.p : nullptr, op.sl.v32.p);
      FH_HIP(hipGetLastError());
    }
    op.sl.filled_version32 = op.version;
    op.sl.scaled32 = jac;
    op.sl.wants_int32 = int32;
  }
}

// after the stream finished: the Jacobi setup's zero-pivot report
void check_pivots(const femhip_solver &s) {
  const int f = (s.prm.preconditioner == FEMHIP_PREC_JACOBI && s.h_flag) ? *s.h_flag : 0;
  FH_CHECK(f != 1, FEMHIP_ERR_ZERO_DIAGONAL, "op1 cannot be entrywise inverted; there exists a null diagonal entry");
  FH_CHECK(f != 2, FEMHIP_ERR_ZERO_DIAGONAL,
           "FP32_REFINED with Jacobi scales by diag(A)^(-1/2): the diagonal must be positive");
}

// Launches a kernel on a stream, or appends it as a kernel node after the
// previous node of a graph under construction.  Graphs are built node by node
// rather than by stream capture: a capture in one thread makes synchronous
// legacy-stream calls of other threads (and the capture itself) fail, and the
// request server sets up and retires requests on several threads at once.
struct Launcher {
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphNode_t last = nullptr;

  template <typename... P, typename... A>
  void operator()(void (*k)(P...), dim3 grid, dim3 block, A... a) {
    std::tuple<std::decay_t<P>...> vals(static_cast<std::decay_t<P>>(a)...);
    auto args = std::apply([](auto &...v) { return std::array<void *, sizeof...(P)>{(void *)&v...}; }, vals);
    if (!graph) {
      FH_HIP(hipLaunchKernel((const void *)k, grid, block, args.data(), 0, stream));
      return;
    }
    hipKernelNodeParams prm{};
    prm.func = (void *)k;
    prm.gridDim = grid;
    prm.blockDim = block;
    prm.sharedMemBytes = 0;
    prm.kernelParams = args.data();
    prm.extra = nullptr;
    hipGraphNode_t node;
    FH_HIP(hipGraphAddKernelNode(&node, graph, last ? &last : nullptr, last ? 1 : 0, &prm));
    last = node;
  }
};

// FP32-refined iteration slots as two kernels (k_rf_a, k_rf_b) on the
// symmetric-diagonal layout, as the three kernels of the FP64 form (k_ru_spmv,
// k_ru_update, k_ru_dir) on the sliced layouts; FEMHIP_RF2=1 / 0 forces two /
// three.  Measured: 128^3 Q1 CG 13.47 ms in two kernels against 13.65 ms in
// three (261 and 263 inner iterations); 64^3 Q2 (class-blocked slices) 95.6 ms
// against 85.9 ms, and slower serving solves, in two.
bool three_kernel_refined(const femhip_solver &s) {
  static const char *env = getenv("FEMHIP_RF2");
  if (env && (atoi(env) == 0 || atoi(env) == 1)) return atoi(env) == 0;
  return !s.op->sl.sym;
}

// Enqueue one iteration of parity `par`: SpMV, update, direction.  With
// `ev`, events are recorded on the launcher's stream around the SpMV; with
// `rh`, the direction kernel also writes the result block there.
void enqueue_one(femhip_solver &s, int par, Launcher &L, const hipEvent_t *ev = nullptr, CgResult *rh = nullptr) {
  femhip_operator &op = *s.op;
  const SlicedDev m = sliced_dev(op);
  const int32_t n = s.n;
  const bool jac = s.prm.preconditioner == FEMHIP_PREC_JACOBI;
  const dim3 B(kThreads), GM(s.g_mat), GV(s.g_vec);
  if (ev) FH_HIP(hipEventRecord(ev[0], L.stream));
  if (!mixed(s)) {
    const double *invd = jac ? s.invd.p : nullptr;
    // reversed odd sweeps: CG 21.35 -> 21.13 ms at 128^3 Q1 and 270.3 -> 268.0 ms at
    // 192^3 on the symmetric copy, but 124.3 -> 125.7 ms at 64^3 Q2 on the
    // class-blocked one, which keeps forward sweeps
    L(k_spmv<double, double>, GM, B, m, op.sl.v64.p, s.p.p, s.q.p, s.st.p, par, s.part_s.p,
      m.alt && m.sym ? par & 1 : 0);
    if (ev) FH_HIP(hipEventRecord(ev[1], L.stream));
    L(k_update<double, double>, GV, B, n, s.r.p, s.q.p, invd, s.st.p, par, s.part_s.p, s.g_mat, s.part_u.p);
    L(k_dir<double>, GV, B, n, s.x.p, s.p.p, s.r.p, invd, s.st.p, par, s.part_u.p, s.g_vec, s.hist.p, s.res.p, rh);
  } else if (!three_kernel_refined(s)) {
    const double *sc = jac ? s.invd.p : nullptr;   // S = diag^(-1/2)
    const SlicedDev m32 = fp32_view(s, m);
    const int rev = m.alt ? par & 1 : 0;
    if (sum32(s))
      L(k_rf_a<float>, GM, B, m, m32, op.sl.v32.p, s.r32.p, s.p32.p, s.q32.p, s.d32.p, s.x.p, sc, s.st.p, par,
        s.part_u.p, s.g_mat, s.part_s.p, s.res.p, rh, rev);
    else
      L(k_rf_a<double>, GM, B, m, m32, op.sl.v32.p, s.r32.p, s.p32.p, s.q32.p, s.d32.p, s.x.p, sc, s.st.p, par,
        s.part_u.p, s.g_mat, s.part_s.p, s.res.p, rh, rev);
    if (ev) FH_HIP(hipEventRecord(ev[1], L.stream));
    L(m.sym ? k_rf_b<true> : k_rf_b<false>, GM, B, m, op.sl.v64.p, s.b.p, s.x.p, s.r32.p, s.q32.p, sc, s.st.p, par, s.part_s.p, s.g_mat,
      s.part_u.p);
  } else {
    const double *sc = jac ? s.invd.p : nullptr;   // S = diag^(-1/2)
    const SlicedDev m32 = fp32_view(s, m);
    const int rev = m.alt ? par & 1 : 0;
    if (sum32(s))
      L(k_ru_spmv<float>, GM, B, m, m32, op.sl.v32.p, op.sl.v64.p, s.p32.p, s.q32.p, s.b.p, s.x.p, s.r32.p, sc,
        s.st.p, par, s.part_s.p, rev);
    else
      L(k_ru_spmv<double>, GM, B, m, m32, op.sl.v32.p, op.sl.v64.p, s.p32.p, s.q32.p, s.b.p, s.x.p, s.r32.p, sc,
        s.st.p, par, s.part_s.p, rev);
    if (ev) FH_HIP(hipEventRecord(ev[1], L.stream));
    L(k_ru_update, GV, B, n, s.r32.p, s.p32.p, s.q32.p, s.st.p, par, s.part_s.p, s.g_mat, s.part_u.p);
    L(k_ru_dir, GV, B, n, s.d32.p, s.p32.p, s.r32.p, s.x.p, sc, s.st.p, par, s.part_u.p, s.g_vec, s.res.p, rh);
  }
}

// Chunks of iterations (even sizes: iteration parity restarts at 0): kChunk
// while convergence is far, kChunkTail near the end, so that few iterations
// are queued past convergence (each still costs its three launches).
constexpr int kChunk = 16, kChunkTail = 4;
constexpr int kChunkSizes[2] = {kChunk, kChunkTail};

// Enqueue a chunk; its last direction kernel writes the result block to `rh`.
void enqueue_iterations(femhip_solver &s, int chunk, Launcher &L, CgResult *rh) {
  for (int i = 0; i < chunk; ++i) enqueue_one(s, i & 1, L, nullptr, i == chunk - 1 ? rh : nullptr);
}

// Chunks run as hipGraphs for small systems, whose iterations are shorter than
// the host's launches; from about a million rows the GPU outpaces the host
// anyway and plain stream launches were measured 1 % faster (128^3: CG 13.64
// against 13.82 ms FP32-refined, 21.19 against 21.31 ms FP64).
// FEMHIP_NO_GRAPH forces stream launches, FEMHIP_GRAPH=1 graphs.
constexpr int32_t kGraphMaxRows = 1 << 20;
bool use_graph(const femhip_solver &s) {
  static const bool no_graph = getenv("FEMHIP_NO_GRAPH") != nullptr;
  static const char *force = getenv("FEMHIP_GRAPH");
  if (no_graph || s.stream_launches) return false;
  if (force && atoi(force) == 1) return true;
  return s.n < kGraphMaxRows;
}

// build (once per precision, chunk size and result slot) the chunk as a hipGraph
void ensure_graph(femhip_solver &s, int size_idx, int slot) {
  const int kind = mixed(s) ? 1 : 0;
  if (!use_graph(s)) return;
  if (s.graph_kind != kind) {
    for (auto &row : s.graphs)
      for (auto &g : row)
        if (g) {
          FH_HIP(hipGraphExecDestroy(g));
          g = nullptr;
        }
    s.graph_kind = kind;
  }
  hipGraphExec_t &g = s.graphs[size_idx][slot];
  if (g) return;
  Launcher L;
  FH_HIP(hipGraphCreate(&L.graph, 0));
  try {
    enqueue_iterations(s, kChunkSizes[size_idx], L, s.d_res_map[slot]);
    FH_HIP(hipGraphInstantiate(&g, L.graph, nullptr, nullptr, 0));
  } catch (...) {
    (void)hipGraphDestroy(L.graph);
    throw;
  }
  FH_HIP(hipGraphDestroy(L.graph));
}

// one chunk; its last kernel writes the result block into pinned slot `slot`
// (mapped host memory), and the slot's event marks the chunk's completion
void push_chunk(femhip_solver &s, int slot, int size_idx = 0) {
  hipStream_t st = s.ctx->stream;
  if (!use_graph(s)) {
    Launcher L;
    L.stream = st;
    enqueue_iterations(s, kChunkSizes[size_idx], L, s.d_res_map[slot]);
  } else {
    ensure_graph(s, size_idx, slot);
    FH_HIP(hipGraphLaunch(s.graphs[size_idx][slot], st));
  }
  FH_HIP(hipEventRecord(s.ev[slot], st));
}

// Launch chunks until the device state says done, keeping one chunk of
// look-ahead.  Near the end the remaining iteration count is extrapolated from
// the stopping quantity at the last two chunk ends (a geometric rate): chunks
// shrink to kChunkTail, and none is added while the one in flight is expected
// to converge.  The plan affects only how many early-exit launches follow
// convergence, never the iterates.
void iterate(femhip_solver &s, int maxit) {
  struct Out {
    int slot, size;
  } q[2];
  int head = 0, nq = 0, next_slot = 0;
  int64_t queued = 0;
  auto push = [&](int size_idx) {
    push_chunk(s, next_slot, size_idx);
    q[(head + nq) & 1] = {next_slot, kChunkSizes[size_idx]};
    ++nq;
    next_slot ^= 1;
    queued += kChunkSizes[size_idx];
  };
  push(0);
  if (queued < maxit) push(0);
  double prev = 0.0;
  int prev_it = -1;
  while (nq > 0) {
    const Out c = q[head];
    head ^= 1;
    --nq;
    FH_HIP(hipEventSynchronize(s.ev[c.slot]));
    const CgResult r = *s.h_res[c.slot];
    if (r.done) break;
    if (queued >= maxit) continue;   // budget queued: wait for the rest
    double need = 1e30;              // iterations still needed
    if (prev_it >= 0 && r.prog_it > prev_it && r.prog > 0.0 && prev > r.prog && r.rhs_test > 0.0) {
      const double rate = std::log(r.prog / prev) / (r.prog_it - prev_it);
      need = r.prog <= r.rhs_test ? 1.0 : std::log(r.rhs_test / r.prog) / rate + 4.0;   // 4: margin
    }
    if (r.prog > 0.0) {
      prev = r.prog;
      prev_it = r.prog_it;
    }
    static const char *plan_env = getenv("FEMHIP_TAIL_CHUNKS");   // 0: always full chunks
    if (plan_env && atoi(plan_env) == 0) need = 1e30;
    const double rest = need - (double)(queued - r.prog_it);   // beyond what is queued
    if (nq > 0 && rest <= 0.0) continue;
    push(rest > kChunk ? 0 : 1);
  }
  FH_HIP(hipStreamSynchronize(s.ctx->stream));
}

void upload_state(femhip_solver &s) {
  CgState h{};
  h.rtol = s.prm.rtol;
  h.atol = s.prm.atol;
  h.maxit = s.prm.max_num_iterations;
  h.crit = s.prm.stopping_criterion;
  h.delta = s.prm.inner_rtol > 0 ? s.prm.inner_rtol : kDefaultDelta;
  FH_HIP(hipMemcpyAsync(s.st.p, &h, sizeof(h), hipMemcpyHostToDevice, s.ctx->stream));
}

// initial residual, direction and device state (x and b already in place)
// A vector in operator order (the RHS or x0; device or host) into solver
// vector dst: copied, or for a permuted layout staged in q (free before a
// solve) and scattered to the solver rows, padding rows cleared.
void load_vector(femhip_solver &s, double *dst, const double *src, hipMemcpyKind kind) {
  hipStream_t st = s.ctx->stream;
  femhip_operator &op = *s.op;
  if (!op.sl.permuted) {
    FH_HIP(hipMemcpyAsync(dst, src, sizeof(double) * (size_t)s.n, kind, st));
    return;
  }
  FH_HIP(hipMemcpyAsync(s.q.p, src, sizeof(double) * (size_t)op.n, kind, st));
  hipLaunchKernelGGL(k_fill0, dim3(1024), dim3(256), 0, st, (int64_t)s.n, dst);
  hipLaunchKernelGGL(k_permute_in, dim3(1024), dim3(256), 0, st, op.n, op.sl.perm.p, s.q.p, dst);
  FH_HIP(hipGetLastError());
}

}  // namespace

// The solution in operator order on the device: x itself, or for a permuted
// layout gathered into q (valid until the next solve).  Synchronizes.
const double *solution_operator_order(const femhip_solver &s) {
  hipStream_t st = s.ctx->stream;
  const femhip_operator &op = *s.op;
  if (op.sl.permuted && op.n > 0) {
    hipLaunchKernelGGL(k_permute_out, dim3(1024), dim3(256), 0, st, op.n, op.sl.perm.p, s.x.p, s.q.p);
    FH_HIP(hipGetLastError());
  }
  FH_HIP(hipStreamSynchronize(st));
  return op.sl.permuted ? s.q.p : s.x.p;
}

namespace {

void begin(femhip_solver &s, bool x_zero) {
  hipStream_t st = s.ctx->stream;
  femhip_operator &op = *s.op;
  const SlicedDev m = sliced_dev(op);
  const bool jac = s.prm.preconditioner == FEMHIP_PREC_JACOBI;
  upload_state(s);
  if (!mixed(s)) {
    hipLaunchKernelGGL((k_init<double>), dim3(s.g_mat), dim3(kThreads), 0, st, m, op.sl.v64.p, s.b.p, s.x.p,
                       x_zero ? 1 : 0, s.r.p, s.p.p, jac ? s.invd.p : nullptr, s.part_i.p);
  } else {
    hipLaunchKernelGGL(k_ru_init, dim3(s.g_mat), dim3(kThreads), 0, st, m, op.sl.v64.p, s.b.p, s.x.p, x_zero ? 1 : 0,
                       s.r.p, s.r32.p, s.p32.p, s.d32.p, jac ? s.invd.p : nullptr, s.part_i.p);
  }
  hipLaunchKernelGGL(mixed(s) ? k_ru_init_fin : k_init_fin, dim3(1), dim3(kThreads), 0, st, s.part_i.p, s.g_mat,
                     s.st.p, s.res.p, (const int *)(jac ? s.d_flag : nullptr));
  FH_HIP(hipGetLastError());
}

// result of the finished iteration (the stream has been synchronized)
void read_result(femhip_solver &s, femhip_cg_result_t *out) {
  CgResult r;
  FH_HIP(hipMemcpy(&r, s.res.p, sizeof(r), hipMemcpyDeviceToHost));
  check_pivots(s);
  out->num_iterations = r.iter;
  out->did_converge = r.converged;
  out->outer_iterations = mixed(s) ? r.updates : 0;
  out->error_estimate = r.est;
  out->rhs_convergence_test = r.rhs_test;
  out->rhs_norm = r.bnorm;
}

void solve(femhip_solver &s, bool x_zero, femhip_cg_result_t *out) {
  begin(s, x_zero);
  iterate(s, s.prm.max_num_iterations);
  read_result(s, out);
}

// One device-resident solve (b = the operator's RHS, x0 = 0) launched kernel
// by kernel with events around every SpMV; returns the mean SpMV duration over
// the iterations that ran (launches queued after convergence return at entry).
void profiled_solve(femhip_solver &s, double *spmv_ms, int32_t *iterations) {
  hipStream_t st = s.ctx->stream;
  load_vector(s, s.b.p, s.op->rhs.p, hipMemcpyDeviceToDevice);
  hipLaunchKernelGGL(k_fill0, dim3(1024), dim3(256), 0, st, (int64_t)s.n, s.x.p);
  begin(s, true);
  Launcher L;
  L.stream = st;
  std::vector<hipEvent_t> ev;
  int launched = 0;
  bool done = false;
  try {
    while (!done && launched < s.prm.max_num_iterations) {
      for (int i = 0; i < kChunk; ++i, ++launched) {
        hipEvent_t e[2];
        FH_HIP(hipEventCreate(&e[0]));
        ev.push_back(e[0]);
        FH_HIP(hipEventCreate(&e[1]));
        ev.push_back(e[1]);
        enqueue_one(s, i & 1, L, e);
      }
      CgResult r;
      FH_HIP(hipMemcpyAsync(&r, s.res.p, sizeof(r), hipMemcpyDeviceToHost, st));
      FH_HIP(hipStreamSynchronize(st));
      done = r.done != 0;
    }
    CgResult r;
    FH_HIP(hipMemcpy(&r, s.res.p, sizeof(r), hipMemcpyDeviceToHost));
    const int it = std::min(r.iter, launched);
    double sum = 0.0;
    for (int k = 0; k < it; ++k) {
      float ms = 0.f;
      FH_HIP(hipEventElapsedTime(&ms, ev[(size_t)2 * k], ev[(size_t)2 * k + 1]));
      sum += ms;
    }
    *spmv_ms = it > 0 ? sum / it : 0.0;
    *iterations = r.iter;
  } catch (...) {
    for (auto e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto e : ev) (void)hipEventDestroy(e);
}

double true_residual(femhip_solver &s) {
  hipStream_t st = s.ctx->stream;
  femhip_operator &op = *s.op;
  hipLaunchKernelGGL(k_true_res, dim3(s.g_mat), dim3(kThreads), 0, st, sliced_dev(op), op.sl.v64.p, s.b.p, s.x.p,
                     s.part_i.p);
  FH_HIP(hipGetLastError());
  std::vector<double> h((size_t)s.g_mat);
  FH_HIP(hipMemcpyAsync(h.data(), s.part_i.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  double sum = 0.0;
  for (double v : h) sum += v;
  return std::sqrt(sum);
}

}  // namespace

// ------------------------------------------- non-blocking driver (server) ---

int solver_chunk() { return kChunk; }

void solver_stream_launches(femhip_solver &s) { s.stream_launches = true; }

void solver_use_host_slots(femhip_solver &s, const SolverHostSlots &hs) {
  FH_CHECK(!s.prepared, FEMHIP_ERR_STATE, "host slots must be set before the solver is prepared");
  for (int k = 0; k < 2; ++k) {
    s.h_res[k] = static_cast<CgResult *>(hs.h_res[k]);
    s.ev[k] = hs.ev[k];
  }
  s.h_flag = hs.h_flag;
  s.borrowed_host = true;
}

size_t solver_result_bytes() { return sizeof(CgResult); }

void solver_start(femhip_solver &s) {
  prepare(s);
  FH_CHECK(s.n > 0, FEMHIP_ERR_STATE, "operator without free DOFs");
  hipStream_t st = s.ctx->stream;
  load_vector(s, s.b.p, s.op->rhs.p, hipMemcpyDeviceToDevice);
  hipLaunchKernelGGL(k_fill0, dim3(1024), dim3(256), 0, st, (int64_t)s.n, s.x.p);
  begin(s, true);
  for (int slot = 0; slot < 2; ++slot) ensure_graph(s, 0, slot);
}

void solver_push(femhip_solver &s, int slot) { push_chunk(s, slot); }

int solver_poll(femhip_solver &s, int slot) {
  const hipError_t e = hipEventQuery(s.ev[slot]);
  if (e == hipErrorNotReady) return 0;
  FH_HIP(e);
  return s.h_res[slot]->done ? 2 : 1;
}

void solver_finish(femhip_solver &s, femhip_cg_result_t *out) {
  FH_HIP(hipStreamSynchronize(s.ctx->stream));
  std::memset(out, 0, sizeof(*out));
  read_result(s, out);
  out->true_residual = true_residual(s);
  out->status = out->did_converge ? FEMHIP_OK : FEMHIP_ERR_NOT_CONVERGED;
  s.have_solution = true;
}

void solver_copy_solution(const femhip_solver &s, double *x) {
  FH_HIP(hipMemcpy(x, solution_operator_order(s), sizeof(double) * (size_t)s.op->n, hipMemcpyDeviceToHost));
}

size_t solver_device_bytes(const femhip_solver &s) {
  size_t b = s.st.bytes() + s.res.bytes();
  for (auto *v : {&s.x, &s.b, &s.r, &s.p, &s.q, &s.invd, &s.part_i, &s.part_s, &s.part_u, &s.part_r, &s.hist})
    b += v->bytes();
  for (auto *v : {&s.r32, &s.p32, &s.q32, &s.d32}) b += v->bytes();
  return b;
}

}  // namespace femhip

extern "C" {

int femhip_solver_create(femhip_ctx_t *ctx, femhip_operator_t *op, const femhip_cg_params_t *params,
                         femhip_solver_t **solver) {
  try {
    FH_CHECK(ctx && op && params && solver, FEMHIP_ERR_INVALID, "null argument");
    const femhip_cg_params_t &p = *params;
    FH_CHECK(p.stopping_criterion >= 0 && p.stopping_criterion <= 9, FEMHIP_ERR_INVALID,
             "unknown stopping criterion");
    FH_CHECK(p.preconditioner == FEMHIP_PREC_IDENTITY || p.preconditioner == FEMHIP_PREC_JACOBI,
             FEMHIP_ERR_INVALID, "unknown preconditioner");
    FH_CHECK(p.precision == FEMHIP_PRECISION_FP64 || p.precision == FEMHIP_PRECISION_FP32_REFINED ||
                 p.precision == FEMHIP_PRECISION_FP32_STRICT,
             FEMHIP_ERR_INVALID, "unknown precision");
    FH_CHECK(p.max_num_iterations >= 1, FEMHIP_ERR_INVALID, "max_num_iterations must be >= 1");
    FH_CHECK(p.rtol >= 0 && p.atol >= 0, FEMHIP_ERR_INVALID, "negative tolerance");
    FH_CHECK(p.precision == FEMHIP_PRECISION_FP64 || p.stopping_criterion == FEMHIP_RES_RES ||
                 p.stopping_criterion == FEMHIP_RES_RHS,
             FEMHIP_ERR_INVALID, "FP32_REFINED supports the RES_RES and RES_RHS criteria");
    FH_CHECK(p.inner_rtol >= 0 && p.inner_rtol < 1, FEMHIP_ERR_INVALID, "reliable-update factor must be in [0,1)");
    FH_HIP(hipSetDevice(ctx->device));
    auto *s = new femhip_solver();
    s->ctx = ctx_retain(ctx);
    op->refs.fetch_add(1);
    s->op = op;
    s->prm = p;
    *solver = s;
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_solver_solve(femhip_solver_t *s, const double *b, double *x, femhip_cg_result_t *res) {
  try {
    FH_CHECK(s && res, FEMHIP_ERR_INVALID, "null argument");
    FH_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    std::memset(res, 0, sizeof(*res));
    hipEvent_t e0, e1;
    FH_HIP(hipEventCreate(&e0));
    FH_HIP(hipEventCreate(&e1));
    FH_HIP(hipEventRecord(e0, st));
    prepare(*s);
    const size_t bytes = sizeof(double) * (size_t)s->op->n;   // host vectors in operator order
    const bool x_zero = (x == nullptr);
    if (s->n > 0) {
      if (b) load_vector(*s, s->b.p, b, hipMemcpyHostToDevice);
      else load_vector(*s, s->b.p, s->op->rhs.p, hipMemcpyDeviceToDevice);
      if (x) load_vector(*s, s->x.p, x, hipMemcpyHostToDevice);
      else hipLaunchKernelGGL(k_fill0, dim3(1024), dim3(256), 0, st, (int64_t)s->n, s->x.p);
      solve(*s, x_zero, res);
    } else {
      res->did_converge = 1;   // no free DOFs: nothing to solve
    }
    FH_HIP(hipEventRecord(e1, st));
    FH_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FH_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    res->solve_ms = ms;
    if (s->n > 0) res->true_residual = true_residual(*s);
    res->status = res->did_converge ? FEMHIP_OK : FEMHIP_ERR_NOT_CONVERGED;
    s->have_solution = true;
    if (x && s->n > 0) FH_HIP(hipMemcpy(x, solution_operator_order(*s), bytes, hipMemcpyDeviceToHost));
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_solver_profile_spmv(femhip_solver_t *s, int32_t reps, double *ms_per_launch, double *bytes) {
  try {
    FH_CHECK(s && reps > 0, FEMHIP_ERR_INVALID, "bad argument");
    FH_HIP(hipSetDevice(s->ctx->device));
    prepare(*s);
    FH_CHECK(s->n > 0, FEMHIP_ERR_STATE, "empty operator");
    femhip_operator &op = *s->op;
    hipStream_t st = s->ctx->stream;
    const bool mx = mixed(*s);
    upload_state(*s);   // done = 0
    hipLaunchKernelGGL(k_fill0, dim3(1024), dim3(256), 0, st, (int64_t)s->n, s->p.p);
    if (mx) hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, st, (int64_t)s->n, s->p.p, s->p32.p);
    const SlicedDev m = sliced_dev(op);
    const SlicedDev m32 = fp32_view(*s, m);
    int i = 0;   // launches alternate their sweep direction as in CG when the layout asks for it
    auto launch = [&] {
      const int rev = m.alt && (mx || m.sym) ? (i++ & 1) : 0;
      if (mx)
        hipLaunchKernelGGL((k_spmv<float, float>), dim3(s->g_mat), dim3(kThreads), 0, st, m32, op.sl.v32.p, s->p32.p,
                           s->q32.p, s->st.p, 0, s->part_s.p, rev);
      else
        hipLaunchKernelGGL((k_spmv<double, double>), dim3(s->g_mat), dim3(kThreads), 0, st, m, op.sl.v64.p, s->p.p,
                           s->q.p, s->st.p, 0, s->part_s.p, rev);
    };
    for (int w = 0; w < 3; ++w) launch();
    hipEvent_t e0, e1;
    FH_HIP(hipEventCreate(&e0));
    FH_HIP(hipEventCreate(&e1));
    FH_HIP(hipEventRecord(e0, st));
    for (int k = 0; k < reps; ++k) launch();
    FH_HIP(hipEventRecord(e1, st));
    FH_HIP(hipEventSynchronize(e1));
    FH_HIP(hipGetLastError());
    float ms = 0.f;
    FH_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_per_launch) *ms_per_launch = ms / reps;
    // algorithmic bytes of the sliced format: values (without the unit
    // diagonal the scaled FP32 copy does not read) + per-slice metadata +
    // diagonal offsets + ELL indices + row lists of sorted windows + one read
    // of p and one write of q
    const double vb = mx ? 4.0 : 8.0;
    const int64_t ell = (int64_t)(op.sl.ecol.n > 1 ? op.sl.ecol.n : 0);
    const double unit = (mx && m32.unit) ? vb * (double)op.sl.stride : 0.0;
    if (bytes)
      *bytes = op.sl.sym ? (double)op.sl.entries * vb - unit + 2.0 * vb * op.n
                         : (double)op.sl.entries * vb + 16.0 * op.sl.ns + 4.0 * (double)op.sl.offs.n +
                               4.0 * (double)ell + 4.0 * kSlice * (double)op.sl.perm_slices + 2.0 * vb * op.n;
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_solver_profile_cg(femhip_solver_t *s, double *spmv_ms, int32_t *iterations, double *bytes) {
  try {
    FH_CHECK(s && spmv_ms && iterations, FEMHIP_ERR_INVALID, "null argument");
    FH_HIP(hipSetDevice(s->ctx->device));
    prepare(*s);
    FH_CHECK(s->n > 0, FEMHIP_ERR_STATE, "empty operator");
    profiled_solve(*s, spmv_ms, iterations);
    s->have_solution = true;
    double b = 0.0, ms = 0.0;
    // algorithmic bytes: the same count as femhip_solver_profile_spmv
    FH_CHECK(femhip_solver_profile_spmv(s, 1, &ms, &b) == FEMHIP_OK, FEMHIP_ERR_HIP, "profile_spmv failed");
    if (bytes) *bytes = b;
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_solver_get_solution(const femhip_solver_t *s, double *x) {
  try {
    FH_CHECK(s && x, FEMHIP_ERR_INVALID, "null argument");
    FH_CHECK(s->have_solution, FEMHIP_ERR_STATE, "no solution yet");
    FH_HIP(hipSetDevice(s->ctx->device));
    if (s->n > 0)
      FH_HIP(hipMemcpy(x, solution_operator_order(*s), sizeof(double) * (size_t)s->op->n, hipMemcpyDeviceToHost));
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_solver_get_convergence_history(const femhip_solver_t *s, int32_t count, double *hist,
                                          int32_t *num_iterations) {
  try {
    FH_CHECK(s && num_iterations && count >= 0 && (hist || count == 0), FEMHIP_ERR_INVALID, "bad argument");
    FH_CHECK(s->have_solution, FEMHIP_ERR_STATE, "no solve yet");
    FH_CHECK(!mixed(*s), FEMHIP_ERR_STATE, "the convergence history is tracked by the FP64 solver");
    FH_HIP(hipSetDevice(s->ctx->device));
    int32_t its = 0;
    if (s->n > 0) {
      CgResult r;
      FH_HIP(hipMemcpy(&r, s->res.p, sizeof(r), hipMemcpyDeviceToHost));
      its = r.iter;
    }
    const int32_t m = std::min(count, its);
    if (m > 0) FH_HIP(hipMemcpy(hist, s->hist.p, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost));
    *num_iterations = its;
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_solver_error_norms(femhip_solver_t *s, int32_t exact, femhip_error_norms_t *norms) {
  try {
    FH_CHECK(s && norms, FEMHIP_ERR_INVALID, "null argument");
    FH_CHECK(s->have_solution, FEMHIP_ERR_STATE, "no solution yet");
    FH_HIP(hipSetDevice(s->ctx->device));
    double out[4] = {0, 0, 0, 0};
    error_norms(*s->op, solution_operator_order(*s), exact, out);
    norms->l2_norm = out[0];
    norms->h1_seminorm = out[1];
    norms->h1_norm = out[2];
    norms->linfty_norm = out[3];
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_solver_destroy(femhip_solver_t *s) {
  if (s) {
    femhip_ctx *c = s->ctx;
    femhip_operator *op = s->op;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete s;
    femhip_operator_destroy(op);
    ctx_release(c);
  }
  return FEMHIP_OK;
}

}  // extern "C"
