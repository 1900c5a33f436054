// norms.hip -- error norms of an FE solution against an exact solution.
//
// FEMPAR's error_norms_scalar_t%compute (Integration/ErrorNorms/
// sbm_error_norms_scalar.i90:67-241): for every cell and quadrature point the
// FE function (sum of nodal values times shape values / physical gradients) is
// subtracted from the exact solution at x_q; detJ w_q |e|^2 and
// detJ w_q |grad e|^2 are summed per cell and over cells (update_norm,
// Integration/ErrorNorms/error_norms.f90:160-172), the max of |e| is kept for
// linfty_norm, and the square roots are taken at the end (finalize_norm,
// :174-221).  One thread per cell; per-block partial sums are reduced on the
// host in block order, so the result is deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "femhip_internal.h"

namespace femhip {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kNormThreads = 256;

struct NormMesh {
  int dim;
  int n[3];
  double lim[6];
};

__device__ __forceinline__ void exact_at(int exact, int dim, const double x[3], double &u, double g[3]) {
  g[0] = g[1] = g[2] = 0.0;
  if (exact == FEMHIP_EXACT_LINEAR) {   // u = x + y (+ z)
    u = 0.0;
    for (int k = 0; k < dim; ++k) {
      u += x[k];
      g[k] = 1.0;
    }
    return;
  }
  double s[3] = {1.0, 1.0, 1.0}, c[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < dim; ++k) {        // u = prod sin(pi x_d)
    s[k] = sin(kPi * x[k]);
    c[k] = cos(kPi * x[k]);
  }
  u = s[0] * s[1] * s[2];
  for (int k = 0; k < dim; ++k) {
    double v = kPi * c[k];
    for (int e = 0; e < dim; ++e)
      if (e != k) v *= s[e];
    g[k] = v;
  }
}

// tab: shp[NQ][ND] | grad[NQ][ND][3] | G[NQ][NV] | dv[NQ]
template <int DIM, int ND, int NQ>
__global__ __launch_bounds__(kNormThreads) void k_error_norms(NormMesh m, int64_t ncells,
                                                               const int32_t *__restrict__ dofs,
                                                               const double *__restrict__ x,
                                                               const double *__restrict__ fixed,
                                                               const double *__restrict__ tab, int exact,
                                                               double *__restrict__ part) {
  constexpr int NV = 1 << DIM;
  const double *shp = tab;
  const double *grad = shp + NQ * ND;
  const double *G = grad + NQ * ND * 3;
  const double *dv = G + NQ * NV;
  double vsum = 0.0, gsum = 0.0, vmax = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * kNormThreads + threadIdx.x; c < ncells;
       c += (int64_t)gridDim.x * kNormThreads) {
    const int ijk[3] = {(int)(c % m.n[0]), (int)((c / m.n[0]) % m.n[1]), (int)(c / ((int64_t)m.n[0] * m.n[1]))};
    double nod[ND];
    for (int i = 0; i < ND; ++i) {
      const int32_t d = dofs[c * ND + i];
      nod[i] = d >= 0 ? x[d] : (fixed ? fixed[-d - 1] : 0.0);
    }
    double vc = 0.0, gc = 0.0;
    for (int q = 0; q < NQ; ++q) {
      double xq[3] = {0.0, 0.0, 0.0};
      for (int v = 0; v < NV; ++v) {   // x_q = sum_v N_v(q) x_v (sbm_cell_map.i90:673-687)
        const double gv = G[q * NV + v];
        for (int k = 0; k < DIM; ++k) {
          const int i = ijk[k] + ((v >> k) & 1);
          const double xv = (m.lim[2 * k + 1] - m.lim[2 * k]) * ((double)i / (double)m.n[k]) + m.lim[2 * k];
          xq[k] += xv * gv;
        }
      }
      double uh = 0.0, gh[3] = {0.0, 0.0, 0.0};
      for (int i = 0; i < ND; ++i) {
        uh += nod[i] * shp[q * ND + i];
        for (int k = 0; k < 3; ++k) gh[k] += nod[i] * grad[(q * ND + i) * 3 + k];
      }
      double u, gu[3];
      exact_at(exact, DIM, xq, u, gu);
      const double e = u - uh;
      double ge2 = 0.0;
      for (int k = 0; k < 3; ++k) ge2 += (gu[k] - gh[k]) * (gu[k] - gh[k]);
      vc += (e * e) * dv[q];
      gc += ge2 * dv[q];
      vmax = fmax(vmax, fabs(e));
    }
    vsum += vc;
    gsum += gc;
  }
  // block reduction: sums and max in a fixed tree order
  __shared__ double sv[kNormThreads], sg[kNormThreads], sm[kNormThreads];
  sv[threadIdx.x] = vsum;
  sg[threadIdx.x] = gsum;
  sm[threadIdx.x] = vmax;
  __syncthreads();
  for (int s = kNormThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sv[threadIdx.x] += sv[threadIdx.x + s];
      sg[threadIdx.x] += sg[threadIdx.x + s];
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 0] = sv[0];
    part[3 * blockIdx.x + 1] = sg[0];
    part[3 * blockIdx.x + 2] = sm[0];
  }
}

}  // namespace

void error_norms(femhip_operator &op, const double *d_x, int exact, double out[4]) {
  femhip_space &s = *op.space;
  const RefFE &fe = s.fe;
  FH_CHECK(fe.ncomp == 1, FEMHIP_ERR_INVALID, "error norms are implemented for scalar spaces");
  FH_CHECK(exact == FEMHIP_EXACT_SIN || exact == FEMHIP_EXACT_LINEAR, FEMHIP_ERR_INVALID, "unknown exact solution");
  if (s.mapped) {   // every cell with its own map
    error_norms_mapped(op, d_x, exact, out);
    return;
  }
  const int nd = fe.nd, nq = fe.nq, nv = 1 << fe.dim;
  std::vector<double> t((size_t)(nq * nd + nq * nd * 3 + nq * nv + nq), 0.0);
  for (int q = 0; q < nq; ++q) {
    for (int i = 0; i < nd; ++i) {
      t[(size_t)(q * nd + i)] = fe.shp[i][q];
      for (int k = 0; k < 3; ++k) t[(size_t)(nq * nd + (q * nd + i) * 3 + k)] = fe.grad[i][k][q];
    }
    for (int v = 0; v < nv; ++v) t[(size_t)(nq * nd * 4 + q * nv + v)] = fe.gshp[v][q];
    t[(size_t)(nq * nd * 4 + nq * nv + q)] = fe.detj[q] * fe.qw[q];
  }
  hipStream_t st = op.ctx->stream;
  DevBuf<double> tab, part;
  tab.alloc(t.size());
  const int64_t blocks64 = std::max<int64_t>(1, std::min<int64_t>(4096, (s.ncells + kNormThreads - 1) / kNormThreads));
  const int blocks = (int)blocks64;
  part.alloc((size_t)3 * blocks);
  FH_HIP(hipMemcpyAsync(tab.p, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, st));
  NormMesh m{};
  m.dim = fe.dim;
  for (int k = 0; k < 3; ++k) m.n[k] = k < fe.dim ? s.mesh.n[k] : 1;
  for (int k = 0; k < 6; ++k) m.lim[k] = s.mesh.limits[k];
  const double *fx = op.fixed.p;   // the values given to the last compute, if any
  const dim3 G(blocks), B(kNormThreads);
  if (fe.dim == 2 && nd == 4)
    hipLaunchKernelGGL((k_error_norms<2, 4, 4>), G, B, 0, st, m, s.ncells, s.d_dofs.p, d_x, fx, tab.p, exact, part.p);
  else if (fe.dim == 2 && nd == 9)
    hipLaunchKernelGGL((k_error_norms<2, 9, 9>), G, B, 0, st, m, s.ncells, s.d_dofs.p, d_x, fx, tab.p, exact, part.p);
  else if (fe.dim == 3 && nd == 8)
    hipLaunchKernelGGL((k_error_norms<3, 8, 8>), G, B, 0, st, m, s.ncells, s.d_dofs.p, d_x, fx, tab.p, exact, part.p);
  else if (fe.dim == 3 && nd == 27)
    hipLaunchKernelGGL((k_error_norms<3, 27, 27>), G, B, 0, st, m, s.ncells, s.d_dofs.p, d_x, fx, tab.p, exact,
                       part.p);
  else
    throw Error{FEMHIP_ERR_INVALID, "unsupported element for error norms"};
  FH_HIP(hipGetLastError());
  std::vector<double> h((size_t)3 * blocks);
  FH_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  double vs = 0.0, gs = 0.0, mx = 0.0;
  for (int b = 0; b < blocks; ++b) {
    vs += h[(size_t)3 * b];
    gs += h[(size_t)3 * b + 1];
    mx = std::max(mx, h[(size_t)3 * b + 2]);
  }
  out[0] = std::sqrt(vs);
  out[1] = std::sqrt(gs);
  out[2] = std::sqrt(vs + gs);
  out[3] = mx;
}

}  // namespace femhip
