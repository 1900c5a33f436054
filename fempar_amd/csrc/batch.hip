// batch.hip -- many independent Poisson problems in one persistent launch.
//
// Problem = FEMPAR's test_poisson setting on its own structured mesh: Q1 on the
// unit square/cube, strong Dirichlet on the whole boundary, constant k, source
// f = f0 k d pi^2 prod sin(pi x) (Sources/Tests/Serial/test_poisson/modules/
// poisson_discrete_integration.f90:62-141 with a coefficient).  With the whole
// boundary fixed, first-touch numbering (Integration/FESpace/
// sbm_serial_fe_space.i90:2623-2922) numbers the free nodes lexicographically,
// so row r is interior node (a,b,c) = r in base m = n-1 and every row's columns
// are r + dx + m dy + m^2 dz, (dx,dy,dz) in {-1,0,1}^d: the matrix is a
// diagonal matrix with 3^d diagonals.  Each problem stores its diagonals
// [k][row] (rows padded to 64) in FP64 and FP32 in shared pools at a
// per-problem offset.
//
// Assembly: one launch over (problem, 256-row chunk) tasks; a thread owns a
// row and sums, over the 2^d cells around its node, the element entries
// coupling it to each neighbour -- every matrix entry is written once by one
// thread (conflict-free, deterministic; the reference's COO duplicate sum in
// owner order).
//
// Solve: one persistent launch; each workgroup takes the next problem from a
// largest-first work queue (single-lane atomic), runs FP32 Jacobi-PCG with
// FP64 reliable updates for it with workgroup-level reductions only, and
// stamps its completion with the device wall clock.  The launch (queue reset +
// solve) is captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "femhip_internal.h"

namespace femhip {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kChunk = 256;   // rows per assembly task

struct BProb {
  int dim, n, m, nrows, K, stride, cls, pad;
  int64_t voff;   // entry offset of diagonal 0 in the value pools
  int64_t xoff;   // entry offset in the vector pools
  double k, f0;
};

struct BOut {
  int iters, updates, conv, pad;
  double rnorm, bnorm;
  long long t_end;
};

struct BTables {             // per (dim, n) class: element stiffness, weighted shapes
  double E[64];              // unit-coefficient element matrix, row-major [i][j]
  double Nw[64];             // detJ w_q N_a(q), [q][a]
  double xi[8][3];           // quadrature points on the reference cell [0,1]^d
  int nd, nq;
};

__device__ __forceinline__ void diag_shift(int k, int dim, int &dx, int &dy, int &dz) {
  dx = k % 3 - 1;
  dy = dim > 1 ? (k / 3) % 3 - 1 : 0;
  dz = dim > 2 ? k / 9 - 1 : 0;
}

// ---------------------------------------------------------------- assembly ---

__global__ __launch_bounds__(kChunk) void k_batch_assemble(const BProb *__restrict__ probs, int count,
                                                           const int64_t *__restrict__ task_ptr,
                                                           const BTables *__restrict__ tabs,
                                                           double *__restrict__ v64, float *__restrict__ v32,
                                                           double *__restrict__ bvec, float *__restrict__ invd) {
  // problem of this block: last p with task_ptr[p] <= blockIdx.x
  int lo = 0, hi = count - 1;
  const int64_t t = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (task_ptr[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const BProb P = probs[lo];
  const BTables &T = tabs[P.cls];
  const int r = (int)(t - task_ptr[lo]) * kChunk + threadIdx.x;
  if (r >= P.stride) return;
  double *V = v64 + P.voff;
  float *V32 = v32 + P.voff;
  if (r >= P.nrows) {   // padding rows
    for (int k = 0; k < P.K; ++k) {
      V[(int64_t)k * P.stride + r] = 0.0;
      V32[(int64_t)k * P.stride + r] = 0.0f;
    }
    bvec[P.xoff + r] = 0.0;
    invd[P.xoff + r] = 0.0f;
    return;
  }
  const int dim = P.dim, m = P.m, nv = 1 << dim;
  int a[3] = {r % m, dim > 1 ? (r / m) % m : 0, dim > 2 ? r / (m * m) : 0};
  int node[3] = {a[0] + 1, a[1] + 1, a[2] + 1};   // vertex index in [1, n-1]
  double diag = 0.0;
  for (int k = 0; k < P.K; ++k) {
    int dd[3];
    diag_shift(k, dim, dd[0], dd[1], dd[2]);
    bool valid = true;
    for (int d = 0; d < dim; ++d) valid = valid && (a[d] + dd[d] >= 0) && (a[d] + dd[d] < m);
    double val = 0.0;
    if (valid) {
      // cells around the node: cell = node - 1 + e (e in {0,1}^d); node is local 1-e
      for (int e = 0; e < nv; ++e) {
        int li = 0, lj = 0;
        bool in = true;
        for (int d = 0; d < dim; ++d) {
          const int ed = (e >> d) & 1;
          const int lid = 1 - ed, ljd = 1 - ed + dd[d];
          in = in && (ljd >= 0) && (ljd <= 1);
          li |= lid << d;
          lj |= (ljd & 1) << d;
        }
        if (in) val += T.E[li * nv + lj];
      }
      val *= P.k;
    }
    V[(int64_t)k * P.stride + r] = val;
    V32[(int64_t)k * P.stride + r] = (float)val;
    // the centre diagonal is k = K/2; with m = 1 other offsets are 0 as well
    // (dx = -m dy), so the offset cannot identify it
    if (k == (P.K >> 1)) diag = val;
  }
  // RHS: sum over the cells around the node of int f N_node
  const double h = 1.0 / P.n;
  double bsum = 0.0;
  for (int e = 0; e < nv; ++e) {
    int li = 0;
    double x0[3];
    for (int d = 0; d < dim; ++d) {
      const int ed = (e >> d) & 1;
      li |= (1 - ed) << d;
      x0[d] = (double)(node[d] - 1 + ed);   // cell origin (vertex index)
    }
    for (int q = 0; q < T.nq; ++q) {
      double u = 1.0;
      for (int d = 0; d < dim; ++d) u *= sin(kPi * ((x0[d] + T.xi[q][d]) * h));
      const double f = P.f0 * P.k * dim * kPi * kPi * u;
      bsum += f * T.Nw[q * nv + li];
    }
  }
  bvec[P.xoff + r] = bsum;
  invd[P.xoff + r] = (float)(1.0 / diag);
}

// ------------------------------------------------------------------- solve ---

template <int BS, int NV>
__device__ __forceinline__ void bsum(double (&v)[NV], double *sm) {
  for (int off = 32; off > 0; off >>= 1)
    for (int j = 0; j < NV; ++j) v[j] += __shfl_down(v[j], off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();   // sm free (previous readers done)
  if (lane == 0)
    for (int j = 0; j < NV; ++j) sm[j * (BS / 64) + w] = v[j];
  __syncthreads();
  for (int j = 0; j < NV; ++j) {
    double s = 0.0;
#pragma unroll 4   // a full unroll at 16 waves holds 16 NV values and spills the 1024-thread solve
    for (int i = 0; i < BS / 64; ++i) s += sm[j * (BS / 64) + i];
    v[j] = s;
  }
}

// Diagonal offsets of a d-dimensional Q1 stencil, k = (dx+1) + 3(dy+1) + 9(dz+1)
template <int DIM>
__device__ __forceinline__ int stencil_off(int k, int m) {
  const int dx = k % 3 - 1;
  const int dy = DIM > 1 ? (k / 3) % 3 - 1 : 0;
  const int dz = DIM > 2 ? k / 9 - 1 : 0;
  return dx + m * (dy + m * dz);
}

// q_i = sum_k V[k][i] p[i + o_k] for this thread's rows (FP32), returns <p, q> partial
template <int DIM>
__device__ __forceinline__ double dia_spmv32(const float *__restrict__ V, int S, int n, int m,
                                             const float *__restrict__ p, float *__restrict__ q, int BS) {
  constexpr int K = DIM == 2 ? 9 : 27;
  int off[K];
#pragma unroll
  for (int k = 0; k < K; ++k) off[k] = stencil_off<DIM>(k, m);
  const int hi = n - 1;
  double pq = 0.0;
  for (int i = threadIdx.x; i < n; i += BS) {
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < K; ++k) acc += V[(int64_t)k * S + i] * p[min(max(i + off[k], 0), hi)];
    q[i] = acc;
    pq += (double)(p[i] * acc);
  }
  return pq;
}

// FP64 residual for a reliable update: r = b - A (x + d); writes r32, returns
// ||r||^2, <r32, M r32>, ||r32||^2 partials
template <int DIM>
__device__ __forceinline__ void dia_residual64(const double *__restrict__ V, int S, int n, int m,
                                               const double *__restrict__ b, const double *__restrict__ x,
                                               const float *__restrict__ d, const float *__restrict__ iv,
                                               float *__restrict__ r, double (&t3)[3], int BS) {
  constexpr int K = DIM == 2 ? 9 : 27;
  int off[K];
#pragma unroll
  for (int k = 0; k < K; ++k) off[k] = stencil_off<DIM>(k, m);
  const int hi = n - 1;
  for (int i = threadIdx.x; i < n; i += BS) {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int c = min(max(i + off[k], 0), hi);
      acc += V[(int64_t)k * S + i] * (x[c] + (double)d[c]);
    }
    const double ri = b[i] - acc;
    const float rf = (float)ri;
    r[i] = rf;
    t3[0] += ri * ri;
    t3[1] += (double)(rf * (iv[i] * rf));
    t3[2] += (double)(rf * rf);
  }
}

struct BPools {
  const double *v64;
  const float *v32;
  const double *b;
  double *x;
  float *r, *p, *q, *d;
  const float *invd;
};

template <int BS>
__device__ void solve_one(const BProb &P, const BPools &pl, double rtol, double delta, int maxit, BOut *out,
                          double *sm) {
  const int n = P.nrows, K = P.K, S = P.stride, hi = n - 1;
  const double *V64 = pl.v64 + P.voff;
  const float *V32 = pl.v32 + P.voff;
  const double *b = pl.b + P.xoff;
  double *x = pl.x + P.xoff;
  float *r = pl.r + P.xoff, *p = pl.p + P.xoff, *q = pl.q + P.xoff, *d = pl.d + P.xoff;
  const float *iv = pl.invd + P.xoff;
  (void)K;
  (void)hi;
  // x0 = 0: r = b, z = M r, p = z, d = 0
  double s3[3] = {0, 0, 0};
  for (int i = threadIdx.x; i < n; i += BS) {
    x[i] = 0.0;
    d[i] = 0.0f;
    const float rf = (float)b[i];
    r[i] = rf;
    const float z = iv[i] * rf;
    p[i] = z;
    s3[0] += (double)(rf * z);
    s3[1] += b[i] * b[i];
    s3[2] += (double)(rf * rf);
  }
  bsum<BS, 3>(s3, sm);
  double rz = s3[0];
  const double bnorm = sqrt(s3[1]);
  const double rhs_test = rtol * bnorm;   // RES_RES with x0 = 0 (cg.f90:305-307)
  double rmax = sqrt(s3[2]), rnorm64 = bnorm;
  // every thread holds the same reduced values; readfirstlane makes the loop
  // and branch conditions wave-uniform for the compiler as well
  int it = 0, upd = 0, conv = __builtin_amdgcn_readfirstlane(bnorm <= rhs_test ? 1 : 0);
  while (!conv && it < maxit) {
    __syncthreads();   // p complete before neighbours are gathered
    double s1[1] = {P.dim == 2 ? dia_spmv32<2>(V32, S, n, P.m, p, q, BS) : dia_spmv32<3>(V32, S, n, P.m, p, q, BS)};
    bsum<BS, 1>(s1, sm);
    const double alpha = s1[0] != 0.0 ? rz / s1[0] : 0.0;
    const float af = (float)alpha;
    double s2[2] = {0, 0};
    float rzl = 0.f, rrl = 0.f;
    for (int i = threadIdx.x; i < n; i += BS) {
      d[i] = d[i] + af * p[i];
      const float ri = r[i] + (-af) * q[i];
      r[i] = ri;
      const float z = iv[i] * ri;
      rzl += ri * z;
      rrl += ri * ri;
    }
    s2[0] = rzl;
    s2[1] = rrl;
    bsum<BS, 2>(s2, sm);
    ++it;
    double rz_new = s2[0];
    const double rnorm = sqrt(s2[1]);
    const int do_ru = __builtin_amdgcn_readfirstlane((rnorm < delta * rmax || rnorm <= rhs_test || it >= maxit) ? 1 : 0);
    if (do_ru) {
      // reliable update: r = b - A64 (x + d) in FP64 replaces the FP32 residual
      __syncthreads();   // d complete before neighbours are gathered
      double t3[3] = {0, 0, 0};
      if (P.dim == 2) dia_residual64<2>(V64, S, n, P.m, b, x, d, iv, r, t3, BS);
      else dia_residual64<3>(V64, S, n, P.m, b, x, d, iv, r, t3, BS);
      bsum<BS, 3>(t3, sm);   // also orders the gathers above before x, d change
      for (int i = threadIdx.x; i < n; i += BS) {
        x[i] = x[i] + (double)d[i];
        d[i] = 0.0f;
      }
      ++upd;
      rnorm64 = sqrt(t3[0]);
      rz_new = t3[1];
      rmax = sqrt(t3[2]);
      if (__builtin_amdgcn_readfirstlane(rnorm64 <= rhs_test ? 1 : 0)) {
        conv = 1;
        break;
      }
      if (it >= maxit) break;
    } else {
      rmax = fmax(rmax, rnorm);
    }
    double beta = 1.0 / rz;
    beta = beta * rz_new;
    rz = rz_new;
    const float bt = (float)beta;
    for (int i = threadIdx.x; i < n; i += BS) {
      const float z = iv[i] * r[i];
      const float t = bt * p[i];
      p[i] = t + z;
    }
  }
  // Wave 0 stores the record (all its lanes store the same values): a branch
  // on threadIdx.x == 0 here, next to the queue fetch of the caller's loop,
  // lets the compiler thread the two lane-0 branches across the loop's back
  // edge and park lane 0 while the other lanes loop on a stale index.
  if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0) {
    out->iters = it;
    out->updates = upd;
    out->conv = conv;
    out->rnorm = rnorm64;
    out->bnorm = bnorm;
    out->t_end = (long long)wall_clock64();
  }
}

__global__ void k_batch_start(int *counter, long long *t0) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *counter = 0;
    *t0 = (long long)wall_clock64();
  }
}

// Persistent solve: workgroups take problems from the queue until it is empty
// (every workgroup reaches the exit once the counter passes `count`).  Every
// branch that encloses or follows a barrier is wave-uniform: the queue fetch is
// made by the whole of wave 0 (lane 0 adds 1, the others 0) and the fetched
// index is read back through readfirstlane.
// BS threads (BS / 64 waves) per problem; 1024 / BS workgroups per CU, each
// held to 128 VGPRs (four waves per SIMD)
template <int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_solve(const BProb *__restrict__ probs, const int *__restrict__ order,
                                                    int count, int *counter, BPools pl, double rtol, double delta,
                                                    int maxit, BOut *out) {
  __shared__ double sm[3 * (BS / 64)];
  __shared__ int s_idx;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (;;) {
    __syncthreads();   // readers of s_idx and sm from the previous problem are done
    if (wave == 0) {
      const int v = atomicAdd(counter, (threadIdx.x & 63) == 0 ? 1 : 0);
      s_idx = __builtin_amdgcn_readfirstlane(v);   // lane 0's fetch
    }
    __syncthreads();
    const int idx = __builtin_amdgcn_readfirstlane(s_idx);
    if (idx >= count) break;
    const int pi = order[idx];
    solve_one<BS>(probs[pi], pl, rtol, delta, maxit, out + pi, sm);
  }
}

}  // namespace
}  // namespace femhip

using namespace femhip;

struct femhip_batch {
  femhip_ctx *ctx = nullptr;
  femhip_cg_params_t prm{};
  std::vector<BProb> probs;
  std::vector<int> order;            // largest first
  std::vector<int64_t> task_ptr;
  int64_t entries = 0, vec = 0;
  DevBuf<BProb> d_probs;
  DevBuf<int> d_order, d_sample, counter;
  DevBuf<int64_t> d_task;
  DevBuf<BTables> d_tabs;
  DevBuf<double> v64, b, x;
  DevBuf<float> v32, r, p, q, d, invd;
  DevBuf<BOut> out;
  DevBuf<long long> t0;
  int threads = 512;
  int wg_per_cu = 2;
  hipGraphExec_t graph = nullptr;
  int graph_threads = 0;
  bool assembled = false;
  ~femhip_batch() {
    if (graph) (void)hipGraphExecDestroy(graph);
  }
};

namespace femhip {
namespace {

BPools pools(femhip_batch &B) {
  return BPools{B.v64.p, B.v32.p, B.b.p, B.x.p, B.r.p, B.p.p, B.q.p, B.d.p, B.invd.p};
}

void launch_solve(femhip_batch &B, const int *order, int count, int threads) {
  hipStream_t st = B.ctx->stream;
  const int grid = std::max(1, B.ctx->num_cus * std::max(1, 1024 / threads));   // 16 waves a CU
  const double delta = B.prm.inner_rtol > 0 ? B.prm.inner_rtol : 0.1;
  hipLaunchKernelGGL(k_batch_start, dim3(1), dim3(64), 0, st, B.counter.p, B.t0.p);
  BPools pl = pools(B);
  switch (threads) {
    case 64:
      hipLaunchKernelGGL(k_batch_solve<64>, dim3(grid), dim3(64), 0, st, B.d_probs.p, order, count, B.counter.p, pl,
                         B.prm.rtol, delta, B.prm.max_num_iterations, B.out.p);
      break;
    case 128:
      hipLaunchKernelGGL(k_batch_solve<128>, dim3(grid), dim3(128), 0, st, B.d_probs.p, order, count, B.counter.p,
                         pl, B.prm.rtol, delta, B.prm.max_num_iterations, B.out.p);
      break;
    case 256:
      hipLaunchKernelGGL(k_batch_solve<256>, dim3(grid), dim3(256), 0, st, B.d_probs.p, order, count, B.counter.p,
                         pl, B.prm.rtol, delta, B.prm.max_num_iterations, B.out.p);
      break;
    case 1024:
      hipLaunchKernelGGL(k_batch_solve<1024>, dim3(grid), dim3(1024), 0, st, B.d_probs.p, order, count,
                         B.counter.p, pl, B.prm.rtol, delta, B.prm.max_num_iterations, B.out.p);
      break;
    default:
      hipLaunchKernelGGL(k_batch_solve<512>, dim3(grid), dim3(512), 0, st, B.d_probs.p, order, count, B.counter.p,
                         pl, B.prm.rtol, delta, B.prm.max_num_iterations, B.out.p);
  }
}

float timed_solve(femhip_batch &B, const int *order, int count, int threads) {
  hipStream_t st = B.ctx->stream;
  hipEvent_t e0, e1;
  FH_HIP(hipEventCreate(&e0));
  FH_HIP(hipEventCreate(&e1));
  FH_HIP(hipEventRecord(e0, st));
  launch_solve(B, order, count, threads);
  FH_HIP(hipGetLastError());
  FH_HIP(hipEventRecord(e1, st));
  FH_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  FH_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

}  // namespace
}  // namespace femhip

extern "C" {

int femhip_batch_create(femhip_ctx_t *ctx, int32_t count, const femhip_batch_problem_t *problems,
                        const femhip_cg_params_t *params, femhip_batch_t **batch) {
  try {
    FH_CHECK(ctx && problems && params && batch && count > 0, FEMHIP_ERR_INVALID, "bad argument");
    FH_CHECK(params->precision == FEMHIP_PRECISION_FP32_REFINED, FEMHIP_ERR_INVALID,
             "the batched solver runs the FP32-refined CG");
    FH_CHECK(params->max_num_iterations >= 1 && params->rtol >= 0, FEMHIP_ERR_INVALID, "bad solver parameters");
    FH_HIP(hipSetDevice(ctx->device));
    auto *B = new femhip_batch();
    B->ctx = ctx_retain(ctx);
    try {
      B->prm = *params;
      // classes (dim, n) -> element tables
      std::vector<std::pair<int, int>> cls;
      std::vector<BTables> tabs;
      B->probs.resize((size_t)count);
      int64_t voff = 0, xoff = 0;
      for (int i = 0; i < count; ++i) {
        const femhip_batch_problem_t &pp = problems[i];
        FH_CHECK((pp.dim == 2 || pp.dim == 3) && pp.n >= 2 && pp.n <= 4096, FEMHIP_ERR_INVALID,
                 "batch problem needs dim 2/3 and 2 <= n <= 4096");
        BProb &P = B->probs[(size_t)i];
        P.dim = pp.dim;
        P.n = pp.n;
        P.m = pp.n - 1;
        P.nrows = 1;
        for (int d = 0; d < pp.dim; ++d) P.nrows *= P.m;
        P.K = pp.dim == 2 ? 9 : 27;
        P.stride = (P.nrows + 63) / 64 * 64;
        P.k = pp.k;
        P.f0 = pp.f0;
        auto key = std::make_pair(pp.dim, pp.n);
        auto it = std::find(cls.begin(), cls.end(), key);
        if (it == cls.end()) {
          RefFE fe;
          const double hh = 1.0 / pp.n;
          double h[3] = {hh, hh, hh};
          build_reffe(fe, pp.dim, 1, 1, h);
          BTables T{};
          T.nd = fe.nd;
          T.nq = fe.nq;
          for (int q = 0; q < fe.nq; ++q) {
            const double factor = fe.detj[q] * fe.qw[q];
            for (int a = 0; a < fe.nd; ++a) {
              for (int b2 = 0; b2 < fe.nd; ++b2) {
                double dot = 0.0;
                for (int k = 0; k < 3; ++k) dot += fe.grad[b2][k][q] * fe.grad[a][k][q];
                T.E[a * fe.nd + b2] += factor * dot;
              }
              T.Nw[q * fe.nd + a] = factor * fe.shp[a][q];
            }
            for (int d = 0; d < 3; ++d) T.xi[q][d] = 0.5 * (fe.qp[d][q] + 1.0);
          }
          cls.push_back(key);
          tabs.push_back(T);
          P.cls = (int)cls.size() - 1;
        } else {
          P.cls = (int)(it - cls.begin());
        }
        P.voff = voff;
        P.xoff = xoff;
        voff += (int64_t)P.K * P.stride;
        xoff += P.stride;
      }
      B->entries = voff;
      B->vec = xoff;
      // largest estimated work first: rows x diagonals x n (CG iterations grow ~ n)
      B->order.resize((size_t)count);
      std::iota(B->order.begin(), B->order.end(), 0);
      std::stable_sort(B->order.begin(), B->order.end(), [&](int a, int b2) {
        const BProb &A = B->probs[(size_t)a], &C = B->probs[(size_t)b2];
        return (double)A.nrows * A.K * A.n > (double)C.nrows * C.K * C.n;
      });
      B->task_ptr.resize((size_t)count + 1);
      B->task_ptr[0] = 0;
      for (int i = 0; i < count; ++i)
        B->task_ptr[(size_t)i + 1] = B->task_ptr[(size_t)i] + (B->probs[(size_t)i].stride + kChunk - 1) / kChunk;
      B->d_probs.alloc((size_t)count);
      B->d_order.alloc((size_t)count);
      B->d_task.alloc((size_t)count + 1);
      B->d_tabs.alloc(tabs.size());
      B->counter.alloc(1);
      B->t0.alloc(1);
      B->out.alloc((size_t)count);
      FH_HIP(hipMemcpy(B->d_probs.p, B->probs.data(), sizeof(BProb) * count, hipMemcpyHostToDevice));
      FH_HIP(hipMemcpy(B->d_order.p, B->order.data(), sizeof(int) * count, hipMemcpyHostToDevice));
      FH_HIP(hipMemcpy(B->d_task.p, B->task_ptr.data(), sizeof(int64_t) * (count + 1), hipMemcpyHostToDevice));
      FH_HIP(hipMemcpy(B->d_tabs.p, tabs.data(), sizeof(BTables) * tabs.size(), hipMemcpyHostToDevice));
      B->v64.alloc((size_t)B->entries);
      B->v32.alloc((size_t)B->entries);
      for (auto *buf : {&B->r, &B->p, &B->q, &B->d, &B->invd}) buf->alloc((size_t)B->vec);
      B->b.alloc((size_t)B->vec);
      B->x.alloc((size_t)B->vec);
      if (const char *t = getenv("FEMHIP_BATCH_THREADS")) {
        const int v = atoi(t);
        FH_CHECK(v == 64 || v == 128 || v == 256 || v == 512 || v == 1024, FEMHIP_ERR_INVALID,
                 "FEMHIP_BATCH_THREADS must be 64, 128, 256, 512 or 1024");
        B->threads = v;
      }
    } catch (...) {
      femhip_batch_destroy(B);
      throw;
    }
    *batch = B;
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_batch_assemble(femhip_batch_t *B, double *ms) {
  try {
    FH_CHECK(B, FEMHIP_ERR_INVALID, "null batch");
    FH_HIP(hipSetDevice(B->ctx->device));
    hipStream_t st = B->ctx->stream;
    hipEvent_t e0, e1;
    FH_HIP(hipEventCreate(&e0));
    FH_HIP(hipEventCreate(&e1));
    FH_HIP(hipEventRecord(e0, st));
    const int64_t tasks = B->task_ptr.back();
    FH_CHECK(tasks < (int64_t)INT32_MAX, FEMHIP_ERR_INVALID, "batch too large");
    hipLaunchKernelGGL(k_batch_assemble, dim3((unsigned)tasks), dim3(kChunk), 0, st, B->d_probs.p,
                       (int)B->probs.size(), B->d_task.p, B->d_tabs.p, B->v64.p, B->v32.p, B->b.p, B->invd.p);
    FH_HIP(hipGetLastError());
    FH_HIP(hipEventRecord(e1, st));
    FH_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    FH_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t;
    B->assembled = true;
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_batch_autotune(femhip_batch_t *B, int32_t *threads) {
  try {
    FH_CHECK(B, FEMHIP_ERR_INVALID, "null batch");
    FH_CHECK(B->assembled, FEMHIP_ERR_STATE, "assemble the batch first");
    FH_HIP(hipSetDevice(B->ctx->device));
    // sample: every k-th problem of the largest-first order, up to 4 per CU
    const int count = (int)B->order.size();
    const int want = std::min(count, 4 * std::max(1, B->ctx->num_cus));
    std::vector<int> sample;
    for (int i = 0; i < want; ++i) sample.push_back(B->order[(size_t)((int64_t)i * count / want)]);
    B->d_sample.alloc(sample.size());
    FH_HIP(hipMemcpy(B->d_sample.p, sample.data(), sizeof(int) * sample.size(), hipMemcpyHostToDevice));
    float best = 1e30f;
    int best_t = B->threads;
    for (int t : {64, 128, 256, 512, 1024}) {   // waves per problem: 1 to 16
      timed_solve(*B, B->d_sample.p, (int)sample.size(), t);   // warm-up
      const float ms = timed_solve(*B, B->d_sample.p, (int)sample.size(), t);
      if (ms < best) {
        best = ms;
        best_t = t;
      }
    }
    B->threads = best_t;
    if (threads) *threads = best_t;
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_batch_solve(femhip_batch_t *B, femhip_batch_result_t *results, double *ms) {
  try {
    FH_CHECK(B, FEMHIP_ERR_INVALID, "null batch");
    FH_CHECK(B->assembled, FEMHIP_ERR_STATE, "assemble the batch first");
    FH_HIP(hipSetDevice(B->ctx->device));
    hipStream_t st = B->ctx->stream;
    if (getenv("FEMHIP_AUTOTUNE") && !B->graph) femhip_batch_autotune(B, nullptr);
    if (!B->graph || B->graph_threads != B->threads) {
      if (B->graph) FH_HIP(hipGraphExecDestroy(B->graph));
      B->graph = nullptr;
      hipGraph_t g;
      FH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      launch_solve(*B, B->d_order.p, (int)B->order.size(), B->threads);
      FH_HIP(hipStreamEndCapture(st, &g));
      FH_HIP(hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0));
      FH_HIP(hipGraphDestroy(g));
      B->graph_threads = B->threads;
    }
    hipEvent_t e0, e1;
    FH_HIP(hipEventCreate(&e0));
    FH_HIP(hipEventCreate(&e1));
    FH_HIP(hipEventRecord(e0, st));
    FH_HIP(hipGraphLaunch(B->graph, st));
    FH_HIP(hipEventRecord(e1, st));
    FH_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    FH_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t;
    if (results) {
      const size_t count = B->probs.size();
      std::vector<BOut> out(count);
      long long t0 = 0;
      FH_HIP(hipMemcpy(out.data(), B->out.p, sizeof(BOut) * count, hipMemcpyDeviceToHost));
      FH_HIP(hipMemcpy(&t0, B->t0.p, sizeof(t0), hipMemcpyDeviceToHost));
      int rate_khz = 0;
      FH_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, B->ctx->device));
      const double us_per_tick = rate_khz > 0 ? 1e3 / rate_khz : 0.01;
      for (size_t i = 0; i < count; ++i) {
        femhip_batch_result_t &R = results[i];
        R.num_iterations = out[i].iters;
        R.updates = out[i].updates;
        R.did_converge = out[i].conv;
        R.num_dofs = B->probs[i].nrows;
        R.true_residual = out[i].rnorm;
        R.rhs_norm = out[i].bnorm;
        R.latency_us = (double)(out[i].t_end - t0) * us_per_tick;
      }
    }
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_batch_get_solution(const femhip_batch_t *B, int32_t i, double *x) {
  try {
    FH_CHECK(B && x && i >= 0 && i < (int)B->probs.size(), FEMHIP_ERR_INVALID, "bad argument");
    FH_HIP(hipSetDevice(B->ctx->device));
    const BProb &P = B->probs[(size_t)i];
    FH_HIP(hipMemcpy(x, B->x.p + P.xoff, sizeof(double) * P.nrows, hipMemcpyDeviceToHost));
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_batch_get_problem_matrix(const femhip_batch_t *B, int32_t i, int32_t *irp, int32_t *ja, double *val,
                                    double *rhs, int64_t *nnz) {
  try {
    FH_CHECK(B && i >= 0 && i < (int)B->probs.size(), FEMHIP_ERR_INVALID, "bad argument");
    FH_CHECK(B->assembled, FEMHIP_ERR_STATE, "assemble the batch first");
    FH_HIP(hipSetDevice(B->ctx->device));
    const BProb &P = B->probs[(size_t)i];
    std::vector<double> V((size_t)P.K * P.stride);
    FH_HIP(hipMemcpy(V.data(), B->v64.p + P.voff, sizeof(double) * V.size(), hipMemcpyDeviceToHost));
    if (rhs) FH_HIP(hipMemcpy(rhs, B->b.p + P.xoff, sizeof(double) * P.nrows, hipMemcpyDeviceToHost));
    int64_t k_out = 0;
    const int m = P.m;
    for (int r = 0; r < P.nrows; ++r) {
      if (irp) irp[r] = (int32_t)(k_out + 1);
      const int a[3] = {r % m, P.dim > 1 ? (r / m) % m : 0, P.dim > 2 ? r / (m * m) : 0};
      for (int k = 0; k < P.K; ++k) {   // ascending offsets = ascending columns
        const int dx = k % 3 - 1, dy = P.dim > 1 ? (k / 3) % 3 - 1 : 0, dz = P.dim > 2 ? k / 9 - 1 : 0;
        const int dd[3] = {dx, dy, dz};
        bool valid = true;
        for (int d = 0; d < P.dim; ++d) valid = valid && a[d] + dd[d] >= 0 && a[d] + dd[d] < m;
        if (!valid) continue;
        const int c = r + dx + m * (dy + m * dz);
        if (ja) ja[k_out] = c + 1;
        if (val) val[k_out] = V[(size_t)k * P.stride + r];
        ++k_out;
      }
    }
    if (irp) irp[P.nrows] = (int32_t)(k_out + 1);
    if (nnz) *nnz = k_out;
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_batch_destroy(femhip_batch_t *B) {
  if (B) {
    femhip_ctx *c = B->ctx;
    if (c) {
      (void)hipSetDevice(c->device);
      (void)hipStreamSynchronize(c->stream);
    }
    delete B;
    ctx_release(c);
  }
  return FEMHIP_OK;
}

}  // extern "C"
