// cg.hip -- (Jacobi-)preconditioned conjugate gradients on gfx950.
//
// Plain path: the reference's cg_solve_body (LinearAlgebra/IterativeLinearSolver/
// cg.f90:101-238) with its stopping criteria (cg.f90:279-407) in FP64.  Three
// kernels per iteration, each a single HBM pass:
//   spmv : q = A p,  <p,q> (and ||p||^2 for the DELTA criteria)  -> alpha
//   upd  : x += alpha p, r -= alpha q, z = invD r, <r,z>, ||r||^2 -> test, beta
//   dir  : p = z + beta p   (z recomputed from r and invD, never stored)
// Scalars live in a device state block.  Each kernel reduces its block partials
// in a fixed order in the last block to finish (deterministic results), and
// every kernel returns at entry once the state says `done`, so the host can
// queue iterations ahead and poll the state only every few iterations.
//
// FP32-refined path: outer FP64 iterative refinement (residual r = b - A x,
// solution update and convergence test in FP64 on the device) around an inner
// FP32 Jacobi-PCG on an FP32 copy of the matrix with FP32 arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "femhip_internal.h"

namespace femhip {
namespace {

constexpr int kThreads = 256;

struct CgState {
  double rz, rnrm_m, alpha, beta;
  double rhs_test, rhs_extra, est, pap, pp;
  double bnorm, r0norm, rnorm;
  double rtol, atol;
  int iter, done, converged, maxit, crit, zero_diag;
};

struct IrState {
  double rnorm, bnorm, r0norm, rhs_test, rtol, atol, inner_rtol;
  int outer, max_outer, done, converged, total_inner, crit;
};

__device__ __forceinline__ bool is_delta(int c) {
  return c == FEMHIP_DELTA_RHS || c == FEMHIP_DELTA_DELTA || c >= FEMHIP_DELTA_RHS_AND_RES_RES;
}
__device__ __forceinline__ bool is_delta_delta(int c) {
  return c == FEMHIP_DELTA_DELTA || c == FEMHIP_DELTA_DELTA_AND_RES_RES || c == FEMHIP_DELTA_DELTA_AND_RES_RHS;
}

// sum NV values over the block; result valid in thread 0
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double sm[NV][kThreads / 64];
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < NV; ++k) v[k] += __shfl_down(v[k], off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < NV; ++k) sm[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int i = 0; i < kThreads / 64; ++i) s += sm[k][i];
      v[k] = s;
    }
}

// write this block's partials; true in every thread of the last block to finish
template <int NV>
__device__ __forceinline__ bool publish_partials(double (&v)[NV], double *part, unsigned *cnt) {
  __shared__ bool last;
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
    for (int k = 0; k < NV; ++k) part[k * gridDim.x + blockIdx.x] = v[k];
    __threadfence();
    last = (atomicAdd(cnt, 1u) == gridDim.x - 1);
  }
  __syncthreads();
  if (last) __threadfence();
  return last;
}

// last block: fixed-order sum of all partials; result valid in thread 0
template <int NV>
__device__ __forceinline__ void gather_partials(double (&v)[NV], const double *part, unsigned *cnt) {
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += kThreads)
    for (int k = 0; k < NV; ++k) v[k] += __hip_atomic_load(part + k * gridDim.x + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  block_sum<NV>(v);
  if (threadIdx.x == 0) *cnt = 0u;
}

// CSR row dot-product by a group of LPR lanes (all lanes of the group call it)
template <int LPR, typename T, typename A>
__device__ __forceinline__ A row_dot(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                     const T *__restrict__ val, const T *__restrict__ x, int32_t r, int lane) {
  A s = 0;
  const int32_t b = rowptr[r], e = rowptr[r + 1];
  for (int32_t k = b + lane; k < e; k += LPR) s += (A)val[k] * (A)x[col[k]];
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_down(s, off, LPR);
  return s;
}

// ------------------------------------------------------------ FP64 / generic CG ---

// r = b - A x ; z = M r ; p = z ; partials <r,z>, <r,r>, <b,b>, <b,M b>
template <int LPR, typename T>
__global__ __launch_bounds__(kThreads) void k_cg_init(int32_t n, const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ col, const T *__restrict__ val,
                                                      const T *__restrict__ b, const T *__restrict__ x, int x_zero,
                                                      T *__restrict__ r, T *__restrict__ p, const T *__restrict__ invd,
                                                      CgState *st, double *part, unsigned *cnt) {
  const int lane = threadIdx.x % LPR;
  const int64_t grp = ((int64_t)blockIdx.x * kThreads + threadIdx.x) / LPR;
  const int64_t ngrp = (int64_t)gridDim.x * kThreads / LPR;
  double v[4] = {0, 0, 0, 0};
  for (int64_t rr = grp; rr < n; rr += ngrp) {
    const int32_t row = (int32_t)rr;
    T ax = x_zero ? (T)0 : row_dot<LPR, T, T>(rowptr, col, val, x, row, lane);
    if (lane == 0) {
      T bi = b[row];
      T ri = bi - ax;
      T zi = invd ? invd[row] * ri : ri;
      r[row] = ri;
      p[row] = zi;
      T mbi = invd ? invd[row] * bi : bi;
      v[0] += (double)(ri * zi);
      v[1] += (double)(ri * ri);
      v[2] += (double)(bi * bi);
      v[3] += (double)(bi * mbi);
    }
  }
  if (!publish_partials<4>(v, part, cnt)) return;
  gather_partials<4>(v, part, cnt);
  if (threadIdx.x == 0) {
    // init_convergence_data (cg.f90:279-321)
    const double rz = v[0], rnorm = sqrt(v[1]), bnorm = sqrt(v[2]), bnrm_m = sqrt(v[3]);
    const double rtol = st->rtol, atol = st->atol;
    const int c = st->crit;
    st->rz = rz;
    st->rnrm_m = sqrt(rz);
    st->bnorm = bnorm;
    st->r0norm = rnorm;
    st->rnorm = rnorm;
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
    st->rhs_test = t;
    st->rhs_extra = extra;
    st->iter = 0;
    st->done = 0;
    st->converged = 0;
  }
}

// q = A p ; <p,q> (, ||p||^2) ; alpha
template <int LPR, typename T, typename A>
__global__ __launch_bounds__(kThreads) void k_cg_spmv(int32_t n, const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ col, const T *__restrict__ val,
                                                      const T *__restrict__ p, T *__restrict__ q, CgState *st,
                                                      double *part, unsigned *cnt) {
  if (st->done) return;
  const int lane = threadIdx.x % LPR;
  const int64_t grp = ((int64_t)blockIdx.x * kThreads + threadIdx.x) / LPR;
  const int64_t ngrp = (int64_t)gridDim.x * kThreads / LPR;
  A pq = 0, pp = 0;
  for (int64_t rr = grp; rr < n; rr += ngrp) {
    const int32_t row = (int32_t)rr;
    A s = row_dot<LPR, T, A>(rowptr, col, val, p, row, lane);
    if (lane == 0) {
      q[row] = (T)s;
      A pi = (A)p[row];
      pq += pi * (A)(T)s;
      pp += pi * pi;
    }
  }
  double v[2] = {(double)pq, (double)pp};
  if (!publish_partials<2>(v, part, cnt)) return;
  gather_partials<2>(v, part, cnt);
  if (threadIdx.x == 0) {
    st->pap = v[0];
    st->pp = v[1];
    st->alpha = (v[0] != 0.0) ? st->rz / v[0] : 0.0;   // cg.f90:195-199
  }
}

// x += alpha p ; r -= alpha q ; <r,z>, ||r||^2 ; stopping test ; beta
template <typename T, typename A>
__global__ __launch_bounds__(kThreads) void k_cg_update(int32_t n, T *__restrict__ x, T *__restrict__ r,
                                                        const T *__restrict__ p, const T *__restrict__ q,
                                                        const T *__restrict__ invd, CgState *st, double *part,
                                                        unsigned *cnt, double *hist) {
  if (st->done) return;
  const T alpha = (T)st->alpha;
  A rz = 0, rr = 0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    x[i] = x[i] + alpha * p[i];
    T ri = r[i] + (-alpha) * q[i];
    r[i] = ri;
    T zi = invd ? invd[i] * ri : ri;
    rz += (A)ri * (A)zi;
    rr += (A)ri * (A)ri;
  }
  double v[2] = {(double)rz, (double)rr};
  if (!publish_partials<2>(v, part, cnt)) return;
  gather_partials<2>(v, part, cnt);
  if (threadIdx.x == 0) {
    // update_convergence_data_and_evaluate_stopping_criteria (cg.f90:323-407)
    const int it = st->iter + 1;
    const int c = st->crit;
    const double rnorm = sqrt(v[1]);
    double est = st->est;
    if (it == 1 && is_delta(c)) {
      est = st->alpha * sqrt(st->pp);
      if (is_delta_delta(c)) st->rhs_test = st->rtol * est + st->atol;
    }
    if (c == FEMHIP_RES_RES || c == FEMHIP_RES_RHS) est = rnorm;
    else if (is_delta(c)) { if (it != 1) est = st->alpha * sqrt(st->pp); }
    else est = st->rnrm_m;
    int conv = est <= st->rhs_test;
    if (c >= FEMHIP_DELTA_RHS_AND_RES_RES && conv) conv = rnorm <= st->rhs_extra;
    st->iter = it;
    st->est = est;
    st->rnorm = rnorm;
    st->converged = conv;
    if (hist) hist[it - 1] = est;
    if (conv || it >= st->maxit) {
      st->done = 1;
    } else {
      double beta = 1.0 / st->rz;
      beta = beta * v[0];
      st->beta = beta;
      st->rz = v[0];
      st->rnrm_m = sqrt(v[0]);
    }
  }
}

// p = z + beta p
template <typename T>
__global__ __launch_bounds__(kThreads) void k_cg_dir(int32_t n, T *__restrict__ p, const T *__restrict__ r,
                                                     const T *__restrict__ invd, const CgState *st) {
  if (st->done) return;
  const T beta = (T)st->beta;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    T zi = invd ? invd[i] * r[i] : r[i];
    T t = beta * p[i];
    p[i] = t + zi;
  }
}

// Jacobi setup: invD = 1 / diag(A) (csr_sparse_matrix.f90:4231-4256, serial_scalar_array.f90:581-588)
__global__ void k_jacobi(int32_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                         const double *__restrict__ val, double *__restrict__ invd, float *__restrict__ invd32,
                         CgState *st) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t lo = rowptr[i], hi = rowptr[i + 1] - 1;
  double d = 0.0;
  while (lo <= hi) {
    int32_t mid = (lo + hi) >> 1;
    int32_t c = col[mid];
    if (c == i) { d = val[mid]; break; }
    if (c < i) lo = mid + 1; else hi = mid - 1;
  }
  if (d == 0.0) st->zero_diag = 1;
  double inv = 1.0 / d;
  invd[i] = inv;
  if (invd32) invd32[i] = (float)inv;
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

__global__ void k_fill0(int64_t n, double *__restrict__ a) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.0;
}

// ---------------------------------------------------------------- refinement ---

// r64 = b - A64 x64 ; ||r||, ||b|| ; outer test ; prepare the inner FP32 rhs
template <int LPR>
__global__ __launch_bounds__(kThreads) void k_ir_residual(int32_t n, const int32_t *__restrict__ rowptr,
                                                          const int32_t *__restrict__ col,
                                                          const double *__restrict__ val, const double *__restrict__ b,
                                                          const double *__restrict__ x, int first, int x_zero,
                                                          double *__restrict__ r, IrState *ir, CgState *in,
                                                          double *part, unsigned *cnt) {
  if (ir->done) return;
  const int lane = threadIdx.x % LPR;
  const int64_t grp = ((int64_t)blockIdx.x * kThreads + threadIdx.x) / LPR;
  const int64_t ngrp = (int64_t)gridDim.x * kThreads / LPR;
  double v[2] = {0, 0};
  for (int64_t rr = grp; rr < n; rr += ngrp) {
    const int32_t row = (int32_t)rr;
    double ax = x_zero ? 0.0 : row_dot<LPR, double, double>(rowptr, col, val, x, row, lane);
    if (lane == 0) {
      double bi = b[row];
      double ri = bi - ax;
      r[row] = ri;
      v[0] += ri * ri;
      v[1] += bi * bi;
    }
  }
  if (!publish_partials<2>(v, part, cnt)) return;
  gather_partials<2>(v, part, cnt);
  if (threadIdx.x == 0) {
    const double rnorm = sqrt(v[0]);
    if (first) {
      ir->bnorm = sqrt(v[1]);
      ir->r0norm = rnorm;
      ir->rhs_test = (ir->crit == FEMHIP_RES_RHS ? ir->bnorm : rnorm) * ir->rtol + ir->atol;
      ir->outer = 0;
      ir->total_inner = 0;
    }
    ir->rnorm = rnorm;
    int conv = rnorm <= ir->rhs_test;
    ir->converged = conv;
    if (conv || ir->outer >= ir->max_outer) ir->done = 1;
    else in->done = 0;
  }
}

// inner start: r32 = r64 / ||r64|| ; z = M r ; p = z ; d = 0 ; inner state init
__global__ __launch_bounds__(kThreads) void k_ir_inner_init(int32_t n, const double *__restrict__ r64,
                                                            float *__restrict__ r32, float *__restrict__ p32,
                                                            float *__restrict__ d32, const float *__restrict__ invd32,
                                                            const IrState *ir, CgState *in, double *part,
                                                            unsigned *cnt) {
  if (ir->done) return;
  const double s = 1.0 / ir->rnorm;
  double v[2] = {0, 0};
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    float ri = (float)(r64[i] * s);
    float zi = invd32 ? invd32[i] * ri : ri;
    r32[i] = ri;
    p32[i] = zi;
    d32[i] = 0.0f;
    v[0] += (double)(ri * zi);
    v[1] += (double)(ri * ri);
  }
  if (!publish_partials<2>(v, part, cnt)) return;
  gather_partials<2>(v, part, cnt);
  if (threadIdx.x == 0) {
    in->rz = v[0];
    in->rnrm_m = sqrt(v[0]);
    in->r0norm = sqrt(v[1]);
    in->rhs_test = ir->inner_rtol * sqrt(v[1]);
    in->iter = 0;
    in->done = 0;
    in->converged = 0;
  }
}

// x64 += ||r64|| * d32 ; bookkeeping of the outer step
__global__ __launch_bounds__(kThreads) void k_ir_update(int32_t n, double *__restrict__ x,
                                                        const float *__restrict__ d32, IrState *ir, CgState *in,
                                                        unsigned *cnt) {
  if (ir->done) return;
  const double s = ir->rnorm;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    x[i] = x[i] + s * (double)d32[i];
  __shared__ bool last;
  if (threadIdx.x == 0) {
    __threadfence();
    last = (atomicAdd(cnt, 1u) == gridDim.x - 1);
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    *cnt = 0u;
    ir->outer += 1;
    ir->total_inner += in->iter;
  }
}

int pick_lpr(double avg_row) {
  if (avg_row <= 6) return 4;
  if (avg_row <= 14) return 8;
  if (avg_row <= 40) return 16;
  return 32;
}

template <typename F>
void dispatch_lpr(int lpr, F &&f) {
  switch (lpr) {
    case 4: f(std::integral_constant<int, 4>()); break;
    case 8: f(std::integral_constant<int, 8>()); break;
    case 16: f(std::integral_constant<int, 16>()); break;
    default: f(std::integral_constant<int, 32>()); break;
  }
}

}  // namespace
}  // namespace femhip

using namespace femhip;

struct femhip_solver {
  femhip_ctx *ctx = nullptr;
  femhip_operator *op = nullptr;
  femhip_cg_params_t prm{};
  int32_t n = 0;
  int nblk = 0, lpr = 8;
  DevBuf<double> x, b, r, p, q, invd, part, hist;
  DevBuf<float> r32, p32, q32, d32, invd32;
  DevBuf<CgState> st, st32;
  DevBuf<IrState> ir;
  DevBuf<unsigned> cnt;
  CgState *h_st = nullptr;   // pinned
  CgState *h_st2 = nullptr;  // pinned, second polling slot
  IrState *h_ir = nullptr;   // pinned
  bool prepared = false;
  bool have_solution = false;
  ~femhip_solver() {
    if (h_st) (void)hipHostFree(h_st);
    if (h_st2) (void)hipHostFree(h_st2);
    if (h_ir) (void)hipHostFree(h_ir);
  }
};

namespace femhip {
namespace {

void prepare(femhip_solver &s) {
  femhip_operator &op = *s.op;
  FH_CHECK(op.computed, FEMHIP_ERR_STATE, "solve before fe_affine_operator%compute");
  hipStream_t st = s.ctx->stream;
  const int32_t n = op.n;
  const bool jac = s.prm.preconditioner == FEMHIP_PREC_JACOBI;
  const bool mixed = s.prm.precision == FEMHIP_PRECISION_FP32_REFINED;
  if (!s.prepared) {
    s.n = n;
    s.x.alloc(n); s.b.alloc(n); s.r.alloc(n); s.p.alloc(n); s.q.alloc(n);
    s.nblk = std::max(1, std::min(2048, (n + kThreads - 1) / kThreads));
    s.part.alloc((size_t)4 * s.nblk);
    s.hist.alloc((size_t)std::max(1, s.prm.max_num_iterations));
    s.st.alloc(1); s.st32.alloc(1); s.ir.alloc(1); s.cnt.alloc(4);
    if (jac) s.invd.alloc(n);
    if (mixed) {
      s.r32.alloc(n); s.p32.alloc(n); s.q32.alloc(n); s.d32.alloc(n);
      if (jac) s.invd32.alloc(n);
    }
    FH_HIP(hipHostMalloc((void **)&s.h_st, sizeof(CgState), hipHostMallocDefault));
    FH_HIP(hipHostMalloc((void **)&s.h_st2, sizeof(CgState), hipHostMallocDefault));
    FH_HIP(hipHostMalloc((void **)&s.h_ir, sizeof(IrState), hipHostMallocDefault));
    FH_HIP(hipMemsetAsync(s.cnt.p, 0, 4 * sizeof(unsigned), st));
    s.lpr = pick_lpr(n ? (double)op.nnz_full / n : 1.0);
    s.prepared = true;
  }
  // matrix-dependent setup (every solve: the operator may have been recomputed)
  CgState h{};
  h.rtol = s.prm.rtol;
  h.atol = s.prm.atol;
  h.maxit = s.prm.max_num_iterations;
  h.crit = s.prm.stopping_criterion;
  FH_HIP(hipMemcpyAsync(s.st.p, &h, sizeof(h), hipMemcpyHostToDevice, st));
  if (jac && n > 0) {
    hipLaunchKernelGGL(k_jacobi, dim3((n + 255) / 256), dim3(256), 0, st, n, op.rowptr.p, op.col.p, op.val.p,
                       s.invd.p, mixed ? s.invd32.p : nullptr, s.st.p);
    FH_HIP(hipGetLastError());
    FH_HIP(hipMemcpyAsync(s.h_st, s.st.p, sizeof(CgState), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    FH_CHECK(!s.h_st->zero_diag, FEMHIP_ERR_ZERO_DIAGONAL,
             "op1 cannot be entrywise inverted; there exists a null diagonal entry");
  }
  if (mixed) {
    op.val32.alloc((size_t)op.nnz_full);
    hipLaunchKernelGGL(k_to_f32, dim3(2048), dim3(256), 0, st, op.nnz_full, op.val.p, op.val32.p);
    FH_HIP(hipGetLastError());
  }
}

void poll(femhip_solver &s, const void *dsrc, void *hdst, size_t bytes, hipEvent_t ev) {
  FH_HIP(hipMemcpyAsync(hdst, dsrc, bytes, hipMemcpyDeviceToHost, s.ctx->stream));
  FH_HIP(hipEventRecord(ev, s.ctx->stream));
}

// run CG iterations on (rowptr, col, val) until the state says done
template <typename T, typename A>
void run_iterations(femhip_solver &s, const T *val, T *x, T *r, T *p, T *q, const T *invd, CgState *dst,
                    double *hist, int maxit) {
  hipStream_t st = s.ctx->stream;
  femhip_operator &op = *s.op;
  const int32_t n = s.n;
  const int nb = s.nblk;
  auto enqueue = [&](int k) {
    for (int i = 0; i < k; ++i) {
      dispatch_lpr(s.lpr, [&](auto L) {
        hipLaunchKernelGGL((k_cg_spmv<decltype(L)::value, T, A>), dim3(nb), dim3(kThreads), 0, st, n, op.rowptr.p,
                           op.col.p, val, p, q, dst, s.part.p, s.cnt.p + 0);
      });
      hipLaunchKernelGGL((k_cg_update<T, A>), dim3(nb), dim3(kThreads), 0, st, n, x, r, p, q, invd, dst, s.part.p,
                         s.cnt.p + 1, hist);
      hipLaunchKernelGGL((k_cg_dir<T>), dim3(nb), dim3(kThreads), 0, st, n, p, r, invd, dst);
    }
  };
  hipEvent_t ev[2];
  FH_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  FH_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  CgState *slot[2] = {s.h_st, s.h_st2};
  const int chunk = 8;
  int queued = 0, cur = 0;
  enqueue(chunk);
  queued += chunk;
  poll(s, dst, slot[cur], sizeof(CgState), ev[cur]);
  for (;;) {
    bool more = queued < maxit;
    if (more) {
      enqueue(chunk);
      queued += chunk;
      poll(s, dst, slot[cur ^ 1], sizeof(CgState), ev[cur ^ 1]);
    }
    FH_HIP(hipEventSynchronize(ev[cur]));
    if (slot[cur]->done || !more) break;
    cur ^= 1;
  }
  FH_HIP(hipStreamSynchronize(st));
  FH_HIP(hipMemcpy(s.h_st, dst, sizeof(CgState), hipMemcpyDeviceToHost));
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
}

double true_residual(femhip_solver &s) {
  // ||b - A x||_2 in FP64 (reporting only)
  hipStream_t st = s.ctx->stream;
  femhip_operator &op = *s.op;
  CgState h{};
  h.crit = FEMHIP_RES_RES;
  FH_HIP(hipMemcpyAsync(s.st32.p, &h, sizeof(h), hipMemcpyHostToDevice, st));
  dispatch_lpr(s.lpr, [&](auto L) {
    hipLaunchKernelGGL((k_cg_init<decltype(L)::value, double>), dim3(s.nblk), dim3(kThreads), 0, st, s.n,
                       op.rowptr.p, op.col.p, op.val.p, s.b.p, s.x.p, 0, s.q.p, s.q.p, (const double *)nullptr,
                       s.st32.p, s.part.p, s.cnt.p + 2);
  });
  FH_HIP(hipGetLastError());
  CgState out{};
  FH_HIP(hipMemcpyAsync(&out, s.st32.p, sizeof(out), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  return out.r0norm;
}

void solve_fp64(femhip_solver &s, bool x_zero, femhip_cg_result_t *res) {
  hipStream_t st = s.ctx->stream;
  femhip_operator &op = *s.op;
  const double *invd = s.prm.preconditioner == FEMHIP_PREC_JACOBI ? s.invd.p : nullptr;
  dispatch_lpr(s.lpr, [&](auto L) {
    hipLaunchKernelGGL((k_cg_init<decltype(L)::value, double>), dim3(s.nblk), dim3(kThreads), 0, st, s.n,
                       op.rowptr.p, op.col.p, op.val.p, s.b.p, s.x.p, x_zero ? 1 : 0, s.r.p, s.p.p, invd, s.st.p,
                       s.part.p, s.cnt.p + 2);
  });
  FH_HIP(hipGetLastError());
  run_iterations<double, double>(s, op.val.p, s.x.p, s.r.p, s.p.p, s.q.p, invd, s.st.p, s.hist.p,
                                 s.prm.max_num_iterations);
  const CgState &h = *s.h_st;
  res->num_iterations = h.iter;
  res->did_converge = h.converged;
  res->outer_iterations = 0;
  res->error_estimate = h.est;
  res->rhs_convergence_test = h.rhs_test;
  res->rhs_norm = h.bnorm;
}

void solve_refined(femhip_solver &s, bool x_zero, femhip_cg_result_t *res) {
  hipStream_t st = s.ctx->stream;
  femhip_operator &op = *s.op;
  const bool jac = s.prm.preconditioner == FEMHIP_PREC_JACOBI;
  IrState hi{};
  hi.rtol = s.prm.rtol;
  hi.atol = s.prm.atol;
  hi.inner_rtol = s.prm.inner_rtol > 0 ? s.prm.inner_rtol : 1e-4;
  hi.max_outer = s.prm.max_outer > 0 ? s.prm.max_outer : 30;
  hi.crit = s.prm.stopping_criterion == FEMHIP_RES_RHS ? FEMHIP_RES_RHS : FEMHIP_RES_RES;
  FH_HIP(hipMemcpyAsync(s.ir.p, &hi, sizeof(hi), hipMemcpyHostToDevice, st));
  CgState h32{};
  h32.maxit = s.prm.max_num_iterations;
  h32.crit = FEMHIP_RES_RES;
  h32.done = 1;
  FH_HIP(hipMemcpyAsync(s.st32.p, &h32, sizeof(h32), hipMemcpyHostToDevice, st));
  if (x_zero) hipLaunchKernelGGL(k_fill0, dim3(1024), dim3(256), 0, st, (int64_t)s.n, s.x.p);
  int first = 1;
  int inner_budget = s.prm.max_num_iterations;
  for (;;) {
    dispatch_lpr(s.lpr, [&](auto L) {
      hipLaunchKernelGGL((k_ir_residual<decltype(L)::value>), dim3(s.nblk), dim3(kThreads), 0, st, s.n,
                         op.rowptr.p, op.col.p, op.val.p, s.b.p, s.x.p, first, first && x_zero ? 1 : 0, s.r.p,
                         s.ir.p, s.st32.p, s.part.p, s.cnt.p + 2);
    });
    FH_HIP(hipGetLastError());
    first = 0;
    FH_HIP(hipMemcpyAsync(s.h_ir, s.ir.p, sizeof(IrState), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    if (s.h_ir->done || inner_budget <= 0) break;
    hipLaunchKernelGGL(k_ir_inner_init, dim3(s.nblk), dim3(kThreads), 0, st, s.n, s.r.p, s.r32.p, s.p32.p, s.d32.p,
                       jac ? s.invd32.p : nullptr, s.ir.p, s.st32.p, s.part.p, s.cnt.p + 3);
    // inner FP32 PCG with its own iteration budget (the remaining total)
    CgState lim{};
    FH_HIP(hipMemcpyAsync(&lim, s.st32.p, sizeof(lim), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    lim.maxit = inner_budget;
    FH_HIP(hipMemcpyAsync(s.st32.p, &lim, sizeof(lim), hipMemcpyHostToDevice, st));
    run_iterations<float, float>(s, op.val32.p, s.d32.p, s.r32.p, s.p32.p, s.q32.p, jac ? s.invd32.p : nullptr,
                                 s.st32.p, nullptr, inner_budget);
    inner_budget -= s.h_st->iter;
    hipLaunchKernelGGL(k_ir_update, dim3(s.nblk), dim3(kThreads), 0, st, s.n, s.x.p, s.d32.p, s.ir.p, s.st32.p,
                       s.cnt.p + 3);
    FH_HIP(hipGetLastError());
  }
  const IrState &h = *s.h_ir;
  res->num_iterations = h.total_inner;
  res->did_converge = h.converged;
  res->outer_iterations = h.outer;
  res->error_estimate = h.rnorm;
  res->rhs_convergence_test = h.rhs_test;
  res->rhs_norm = h.bnorm;
}

}  // namespace
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
    FH_CHECK(p.precision == FEMHIP_PRECISION_FP64 || p.precision == FEMHIP_PRECISION_FP32_REFINED,
             FEMHIP_ERR_INVALID, "unknown precision");
    FH_CHECK(p.max_num_iterations >= 1, FEMHIP_ERR_INVALID, "max_num_iterations must be >= 1");
    FH_CHECK(p.rtol >= 0 && p.atol >= 0, FEMHIP_ERR_INVALID, "negative tolerance");
    FH_CHECK(p.precision == FEMHIP_PRECISION_FP64 || p.stopping_criterion == FEMHIP_RES_RES ||
                 p.stopping_criterion == FEMHIP_RES_RHS,
             FEMHIP_ERR_INVALID, "FP32_REFINED supports the RES_RES and RES_RHS criteria");
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
    hipEvent_t e0, e1;
    FH_HIP(hipEventCreate(&e0));
    FH_HIP(hipEventCreate(&e1));
    FH_HIP(hipEventRecord(e0, st));
    prepare(*s);
    const size_t bytes = sizeof(double) * (size_t)s->n;
    if (b) FH_HIP(hipMemcpyAsync(s->b.p, b, bytes, hipMemcpyHostToDevice, st));
    else FH_HIP(hipMemcpyAsync(s->b.p, s->op->rhs.p, bytes, hipMemcpyDeviceToDevice, st));
    bool x_zero = (x == nullptr);
    if (x) FH_HIP(hipMemcpyAsync(s->x.p, x, bytes, hipMemcpyHostToDevice, st));
    else hipLaunchKernelGGL(k_fill0, dim3(1024), dim3(256), 0, st, (int64_t)s->n, s->x.p);
    std::memset(res, 0, sizeof(*res));
    if (s->n == 0) {   // no free DOFs: nothing to solve
      res->did_converge = 1;
      FH_HIP(hipEventRecord(e1, st));
      FH_HIP(hipEventSynchronize(e1));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      s->have_solution = true;
      return FEMHIP_OK;
    }
    if (s->prm.precision == FEMHIP_PRECISION_FP64) solve_fp64(*s, x_zero, res);
    else solve_refined(*s, x_zero, res);
    FH_HIP(hipEventRecord(e1, st));
    FH_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FH_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    res->solve_ms = ms;
    res->true_residual = true_residual(*s);
    res->status = res->did_converge ? FEMHIP_OK : FEMHIP_ERR_NOT_CONVERGED;
    s->have_solution = true;
    if (x) FH_HIP(hipMemcpy(x, s->x.p, bytes, hipMemcpyDeviceToHost));
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
    femhip_operator &op = *s->op;
    hipStream_t st = s->ctx->stream;
    const bool mixed = s->prm.precision == FEMHIP_PRECISION_FP32_REFINED;
    CgState h{};
    h.rz = 1.0;
    FH_HIP(hipMemcpyAsync(s->st32.p, &h, sizeof(h), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fill0, dim3(1024), dim3(256), 0, st, (int64_t)s->n, s->p.p);
    if (mixed) hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, st, (int64_t)s->n, s->p.p, s->p32.p);
    auto launch = [&] {
      dispatch_lpr(s->lpr, [&](auto L) {
        if (mixed)
          hipLaunchKernelGGL((k_cg_spmv<decltype(L)::value, float, float>), dim3(s->nblk), dim3(kThreads), 0, st,
                             s->n, op.rowptr.p, op.col.p, op.val32.p, s->p32.p, s->q32.p, s->st32.p, s->part.p,
                             s->cnt.p + 0);
        else
          hipLaunchKernelGGL((k_cg_spmv<decltype(L)::value, double, double>), dim3(s->nblk), dim3(kThreads), 0, st,
                             s->n, op.rowptr.p, op.col.p, op.val.p, s->p.p, s->q.p, s->st32.p, s->part.p,
                             s->cnt.p + 0);
      });
    };
    for (int i = 0; i < 3; ++i) launch();
    hipEvent_t e0, e1;
    FH_HIP(hipEventCreate(&e0));
    FH_HIP(hipEventCreate(&e1));
    FH_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < reps; ++i) launch();
    FH_HIP(hipEventRecord(e1, st));
    FH_HIP(hipEventSynchronize(e1));
    FH_HIP(hipGetLastError());
    float ms = 0.f;
    FH_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_per_launch) *ms_per_launch = ms / reps;
    const double vb = mixed ? 4.0 : 8.0;
    if (bytes) *bytes = (double)op.nnz_full * (vb + 4.0) + 4.0 * (op.n + 1) + 2.0 * vb * op.n;
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
    FH_HIP(hipMemcpy(x, s->x.p, sizeof(double) * (size_t)s->n, hipMemcpyDeviceToHost));
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
