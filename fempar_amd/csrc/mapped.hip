// mapped.hip -- cell-wise integration for cells that are not translates of
// one another, and for coefficients given at the quadrature points.
//
// When the caller moves the vertices of the structured mesh
// (femhip_space_set_vertex_coords), every cell has its own trilinear map and
// the reference recomputes it per cell (cell_ressemblance = no_ressemblance,
// Integration/FESpace/sbm_fe_cell_iterator.i90:970-987): the Jacobian at each
// quadrature point is the linear combination of the geometric Q1 gradients and
// the cell's vertex coordinates (Integration/ReferenceFE/sbm_cell_map.i90:
// 245-369, field.f90:763-787), inverted with invmtx
// (sbm_cell_integrator.i90:448-560); physical gradients are
// invJ^T * reference gradient (sbm_lagrangian_reference_fe.i90:1334-1341), and
// the quadrature points' coordinates are the geometric map of the reference
// points (sbm_cell_map.i90:673-687).  The element integrands are those of the
// plugins: Poisson (Sources/Tests/Serial/test_poisson/modules/
// poisson_discrete_integration.f90:109-130, with the coefficient at the
// quadrature point) and irreducible elasticity
// (sbm_irreducible_discrete_integration.i90:64-130).
//
// Work split: k_elmat_mapped forms every cell's element matrix and vector (one
// wavefront per cell: the map at each point on its own lane, the physical
// gradients of all (shape, point) pairs over the lanes in LDS, then one entry
// per lane and step) into a cells x nd x nd array in HBM; the rows' owners
// then gather them in ascending cell order (k_gather_q2 for 3-D Q2,
// k_assemble_owner otherwise), so the result is deterministic, written once
// per value, and bitwise symmetric (entry (i, j) and (j, i) are the same
// products summed in the same order).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "femhip_internal.h"

namespace femhip {
namespace {

constexpr double kPi = 3.14159265358979323846;

struct MapMesh {
  int dim;
  int n[3];       // cells per direction (1 beyond dim)
  int nv[3];      // vertices per direction (1 beyond dim)
  double lo[3], hi[3];
};

struct MapPhys {
  int kind, k_kind, f_kind;
  double k0, f0;
  double lambda, mu;
  double body[3];
};

MapMesh map_mesh(const femhip_space &s) {
  MapMesh m{};
  m.dim = s.mesh.dim;
  for (int k = 0; k < 3; ++k) {
    m.n[k] = k < m.dim ? s.mesh.n[k] : 1;
    m.nv[k] = k < m.dim ? s.mesh.n[k] + 1 : 1;
    m.lo[k] = k < m.dim ? s.mesh.limits[2 * k] : 0.0;
    m.hi[k] = k < m.dim ? s.mesh.limits[2 * k + 1] : 0.0;
  }
  return m;
}

// Vertex a (bit k: +1 along axis k) of cell c: moved coordinates, or the
// structured generator's (uniform_hex_mesh_generator.f90:608-612, 633-635).
template <int DIM>
__device__ __forceinline__ void cell_vertex(const MapMesh &m, const double *__restrict__ vc, int64_t c, int a,
                                            double x[3]) {
  const int i0 = (int)(c % m.n[0]), i1 = (int)((c / m.n[0]) % m.n[1]), i2 = (int)(c / ((int64_t)m.n[0] * m.n[1]));
  const int ijk[3] = {i0 + (a & 1), i1 + ((a >> 1) & 1), i2 + ((a >> 2) & 1)};
  x[0] = x[1] = x[2] = 0.0;
  if (vc) {
    const int64_t v = ijk[0] + (int64_t)m.nv[0] * (ijk[1] + (int64_t)m.nv[1] * ijk[2]);
#pragma unroll
    for (int k = 0; k < DIM; ++k) x[k] = vc[v * DIM + k];
  } else {
#pragma unroll
    for (int k = 0; k < DIM; ++k) x[k] = (m.hi[k] - m.lo[k]) * ((double)ijk[k] / (double)m.n[k]) + m.lo[k];
  }
}

// Jacobian J[l][j] = sum_a dN_a/dxi_j x_a[l] at point q, its inverse B and
// determinant, in the reference's loop and formula order (sbm_cell_map.i90,
// invmtx of sbm_cell_integrator.i90:448-560).
template <int DIM, int NQ>
__device__ __forceinline__ double jacobian(const double (*xv)[3], const double *__restrict__ gdshp, int q,
                                           double B[3][3]) {
  constexpr int NV = 1 << DIM;
  double J[3][3];
  for (int j = 0; j < 3; ++j) {
    for (int l = 0; l < 3; ++l) J[l][j] = 0.0;
    for (int a = 0; a < NV; ++a) {
      const double alpha = gdshp[(a * 3 + j) * NQ + q];
      for (int l = 0; l < 3; ++l) J[l][j] = J[l][j] + alpha * xv[a][l];
    }
  }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) B[a][b] = a == b ? 1.0 : 0.0;
  double det;
  if (DIM == 2) {
    det = J[0][0] * J[1][1] - J[1][0] * J[0][1];
    const double den = 1.0 / det;
    B[0][0] = J[1][1] * den;
    B[1][1] = J[0][0] * den;
    B[1][0] = -J[1][0] * den;
    B[0][1] = -J[0][1] * den;
  } else {
    const double t1 = J[1][1] * J[2][2] - J[2][1] * J[1][2];
    const double t2 = -J[1][0] * J[2][2] + J[2][0] * J[1][2];
    const double t3 = J[1][0] * J[2][1] - J[2][0] * J[1][1];
    det = J[0][0] * t1 + J[0][1] * t2 + J[0][2] * t3;
    const double den = 1.0 / det;
    B[0][0] = t1 * den;
    B[1][0] = t2 * den;
    B[2][0] = t3 * den;
    B[1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) * den;
    B[2][1] = (-J[0][0] * J[2][1] + J[0][1] * J[2][0]) * den;
    B[2][2] = (J[0][0] * J[1][1] - J[1][0] * J[0][1]) * den;
    B[0][1] = (-J[0][1] * J[2][2] + J[2][1] * J[0][2]) * den;
    B[0][2] = (J[0][1] * J[1][2] - J[1][1] * J[0][2]) * den;
    B[1][2] = (-J[0][0] * J[1][2] + J[1][0] * J[0][2]) * den;
  }
  return det;
}

// x_q = sum_a N_a(q) x_a (sbm_cell_map.i90:673-687)
template <int DIM, int NQ>
__device__ __forceinline__ void qpoint(const double (*xv)[3], const double *__restrict__ gshp, int q, double x[3]) {
  constexpr int NV = 1 << DIM;
  x[0] = x[1] = x[2] = 0.0;
  for (int a = 0; a < NV; ++a)
    for (int k = 0; k < 3; ++k) x[k] = x[k] + xv[a][k] * gshp[a * NQ + q];
}

// coefficient and source at a quadrature point (the kinds of femhip.h)
template <int DIM>
__device__ __forceinline__ double k_at(const MapPhys &ph, const double x[3], int64_t cell, int q, int nq,
                                       const double *kcell, const double *kqp) {
  switch (ph.k_kind) {
    case FEMHIP_K_SIN3:
      return 1.0 + 0.5 * sin(2 * kPi * x[0]) * sin(2 * kPi * x[1]) * (DIM == 3 ? sin(2 * kPi * x[2]) : 1.0);
    case FEMHIP_K_CELL: return kcell[cell];
    case FEMHIP_K_QPOINT: return kqp[cell * nq + q];
    default: return ph.k0;
  }
}

template <int DIM>
__device__ __forceinline__ double f_at(const MapPhys &ph, const double x[3], double kk, int64_t cell, int q, int nq,
                                       const double *fqp) {
  switch (ph.f_kind) {
    case FEMHIP_F_MANUFACTURED: {   // f = -div(k grad u) = k d pi^2 u - grad k . grad u, u = prod sin(pi x_k)
      double su[3], cu[3];
      for (int k = 0; k < 3; ++k) {
        su[k] = k < DIM ? sin(kPi * x[k]) : 1.0;
        cu[k] = k < DIM ? cos(kPi * x[k]) : 0.0;
      }
      const double u = su[0] * su[1] * su[2];
      double f = kk * DIM * kPi * kPi * u;
      if (ph.k_kind == FEMHIP_K_SIN3) {
        double s2[3], c2[3];
        for (int k = 0; k < 3; ++k) {
          s2[k] = k < DIM ? sin(2 * kPi * x[k]) : 1.0;
          c2[k] = k < DIM ? cos(2 * kPi * x[k]) : 0.0;
        }
        const double gk[3] = {kPi * c2[0] * s2[1] * s2[2], kPi * s2[0] * c2[1] * s2[2], kPi * s2[0] * s2[1] * c2[2]};
        const double gu[3] = {kPi * cu[0] * su[1] * su[2], kPi * su[0] * cu[1] * su[2], kPi * su[0] * su[1] * cu[2]};
        double dot = 0.0;
        for (int k = 0; k < DIM; ++k) dot += gk[k] * gu[k];
        f -= dot;
      }
      return f;
    }
    case FEMHIP_F_CELL_SIN: {
      double u = 1.0;
      for (int k = 0; k < DIM; ++k) u *= sin(kPi * x[k]);
      return ph.f0 * kk * DIM * kPi * kPi * u;
    }
    case FEMHIP_F_QPOINT: return fqp[cell * nq + q];
    default: return ph.f0;
  }
}

// Table of the reference FE (mapped_table): qw[NQ] | shp[NN][NQ] | dshp[NN][3][NQ]
// | gshp[8][NQ] | gdshp[8][3][NQ]
struct TabView {
  const double *qw, *shp, *dshp, *gshp, *gdshp;
};
template <int NN, int NQ>
__device__ __forceinline__ TabView tab_view(const double *t) {
  TabView v;
  v.qw = t;
  v.shp = v.qw + NQ;
  v.dshp = v.shp + NN * NQ;
  v.gshp = v.dshp + NN * 3 * NQ;
  v.gdshp = v.gshp + 8 * NQ;
  return v;
}

// PHYS 0: Poisson, elmat(i, j) = sum_q (detJ w_q k_q) (grad phi_j . grad phi_i),
// elvec(i) = sum_q (detJ w_q f_q) phi_i (poisson_discrete_integration.f90:116-128)
// PHYS 1: elasticity, vector Q1 with component-major local DOFs (i = c nn + a),
// elmat(i, j) = sum_q dV (lambda div div + 2 mu eps : eps), elvec(i) = sum_q dV b_c phi_a
// Output row-major: elmat[cell][i][j] (i the row's local DOF), elvec[cell][i].
template <int DIM, int NN, int NQ, int PHYS>
__global__ __launch_bounds__(64) void k_elmat_mapped(MapPhys ph, MapMesh m, int64_t ncells,
                                                     const double *__restrict__ tab, const double *__restrict__ vc,
                                                     const double *__restrict__ kcell, const double *__restrict__ kqp,
                                                     const double *__restrict__ fqp, double *__restrict__ elmat,
                                                     double *__restrict__ elvec) {
  constexpr int NV = 1 << DIM;
  constexpr int ND = PHYS == 0 ? NN : NN * DIM;
  const TabView tv = tab_view<NN, NQ>(tab);
  __shared__ double sx[NV][3];
  __shared__ double sfac[NQ], sk[NQ], sf[NQ];
  __shared__ double sB[NQ][3][3];
  __shared__ double sg[NN][3][NQ];
  const int64_t cell = blockIdx.x;
  const int lane = threadIdx.x;
  if (lane < NV) cell_vertex<DIM>(m, vc, cell, lane, sx[lane]);
  __syncthreads();
  if (lane < NQ) {
    const int q = lane;
    double B[3][3];
    const double det = jacobian<DIM, NQ>(sx, tv.gdshp, q, B);
    sfac[q] = fabs(det) * tv.qw[q];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) sB[q][a][b] = B[a][b];
    if (PHYS == 0) {
      double x[3];
      qpoint<DIM, NQ>(sx, tv.gshp, q, x);
      const double kk = k_at<DIM>(ph, x, cell, q, NQ, kcell, kqp);
      sk[q] = kk;
      sf[q] = f_at<DIM>(ph, x, kk, cell, q, NQ, fqp);
    }
  }
  __syncthreads();
  for (int p = lane; p < NN * NQ; p += 64) {   // physical gradients: phy(i) += invJ(j,i) ref(j)
    const int a = p / NQ, q = p % NQ;
    for (int i = 0; i < 3; ++i) {
      double g = 0.0;
      for (int j = 0; j < 3; ++j) g = g + sB[q][j][i] * tv.dshp[(a * 3 + j) * NQ + q];
      sg[a][i][q] = g;
    }
  }
  __syncthreads();
  double *em = elmat + cell * (ND * ND);
  for (int e = lane; e < ND * ND; e += 64) {
    const int i = e / ND, j = e % ND;
    double acc = 0.0;
    if (PHYS == 0) {
      for (int q = 0; q < NQ; ++q) {
        double dot = 0.0;
        for (int k = 0; k < 3; ++k) dot += sg[j][k][q] * sg[i][k][q];
        acc = acc + sfac[q] * sk[q] * dot;
      }
    } else {
      const int ci = i / NN, ai = i % NN, cj = j / NN, aj = j % NN;
      for (int q = 0; q < NQ; ++q) {
        const double gi[3] = {sg[ai][0][q], sg[ai][1][q], sg[ai][2][q]};
        const double gj[3] = {sg[aj][0][q], sg[aj][1][q], sg[aj][2][q]};
        const double divdiv = gi[ci] * gj[cj];
        const double epseps =
            0.5 * ((ci == cj ? gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2] : 0.0) + gi[cj] * gj[ci]);
        acc += sfac[q] * (ph.lambda * divdiv + 2 * ph.mu * epseps);
      }
    }
    em[e] = acc;
  }
  if (lane < ND) {
    double be = 0.0;
    if (PHYS == 0) {
      for (int q = 0; q < NQ; ++q) be = be + sfac[q] * sf[q] * tv.shp[lane * NQ + q];
    } else {
      const int c = lane / NN, a = lane % NN;
      for (int q = 0; q < NQ; ++q) be += sfac[q] * ph.body[c] * tv.shp[a * NQ + q];
    }
    elvec[cell * ND + lane] = be;
  }
}

// coordinates of every quadrature point: xq[(cell * NQ + q) * DIM + k]
template <int DIM, int NQ>
__global__ void k_qpoints(MapMesh m, int64_t ncells, const double *__restrict__ tab, const double *__restrict__ vc,
                          double *__restrict__ xq) {
  constexpr int NV = 1 << DIM;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const double *gshp = tab + NQ;   // tab = gshp table only: [NV][NQ] after a leading qw[NQ]
  double xv[NV][3];
  for (int a = 0; a < NV; ++a) cell_vertex<DIM>(m, vc, c, a, xv[a]);
  for (int q = 0; q < NQ; ++q) {
    double x[3];
    qpoint<DIM, NQ>(xv, gshp, q, x);
    for (int k = 0; k < DIM; ++k) xq[(c * NQ + q) * DIM + k] = x[k];
  }
}

// g = x + y (+ z) at each fixed DOF's node, the node being the trilinear map of
// its reference position xi = (2 / order) m - 1 (sbm_serial_fe_space.i90:
// 997-1176); taken from the last cell touching the node, as the reference's
// cell loop leaves the last write
template <int DIM>
__global__ void k_interp_mapped(FaceTab t, MapMesh m, int64_t ncells, const int32_t *__restrict__ dofs,
                                const double *__restrict__ vc, double *__restrict__ out) {
#pragma clang fp contract(off)
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  int ijk[3];
  cell_ijk(t, c, ijk);
  constexpr int nv = 1 << DIM;   // the vertex coordinates stay in registers
  double xv[nv][3];
#pragma unroll
  for (int a = 0; a < nv; ++a) cell_vertex<DIM>(m, vc, c, a, xv[a]);
  for (int l = 0; l < t.nd; ++l) {
    const int32_t d = dofs[c * t.nd + l];
    if (d >= 0) continue;
    const int f = t.face_of[l];
    bool last = true;
    for (int k = 0; k < DIM; ++k) {
      if ((t.type[f] >> k) & 1) continue;
      const int w = ijk[k] + ((t.anchor[f] >> k) & 1);
      last = last && ijk[k] == min(w, t.n[k] - 1);
    }
    if (!last) continue;
    int rem = l % t.nnode;
    double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const int mm = rem % t.np1;
      rem /= t.np1;
      xi[k] = (2.0 / t.order) * mm - 1.0;
    }
    double x[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < nv; ++a) {
      double N = 1.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) N *= ((a >> k) & 1) ? 0.5 * (1.0 + xi[k]) : 0.5 * (1.0 - xi[k]);
#pragma unroll
      for (int k = 0; k < 3; ++k) x[k] += xv[a][k] * N;
    }
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < DIM; ++k) v += x[k];
    out[-d - 1] = v;
  }
}

// error norms with each cell's own map (sbm_error_norms_scalar.i90:67-241):
// per-block partial sums of detJ w |e|^2, detJ w |grad e|^2 and max |e|
constexpr int kNormThreads = 256;
template <int DIM, int NN, int NQ>
__global__ __launch_bounds__(kNormThreads) void k_error_norms_mapped(MapMesh m, int64_t ncells,
                                                                      const int32_t *__restrict__ dofs,
                                                                      const double *__restrict__ x,
                                                                      const double *__restrict__ fixed,
                                                                      const double *__restrict__ tab,
                                                                      const double *__restrict__ vc, int exact,
                                                                      double *__restrict__ part) {
  constexpr int NV = 1 << DIM;
  const TabView tv = tab_view<NN, NQ>(tab);
  double vsum = 0.0, gsum = 0.0, vmax = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * kNormThreads + threadIdx.x; c < ncells;
       c += (int64_t)gridDim.x * kNormThreads) {
    double xv[NV][3];
    for (int a = 0; a < NV; ++a) cell_vertex<DIM>(m, vc, c, a, xv[a]);
    double nod[NN];
    for (int i = 0; i < NN; ++i) {
      const int32_t d = dofs[c * NN + i];
      nod[i] = d >= 0 ? x[d] : (fixed ? fixed[-d - 1] : 0.0);
    }
    double vc2 = 0.0, gc2 = 0.0;
    for (int q = 0; q < NQ; ++q) {
      double B[3][3];
      const double dv = fabs(jacobian<DIM, NQ>(xv, tv.gdshp, q, B)) * tv.qw[q];
      double xq[3];
      qpoint<DIM, NQ>(xv, tv.gshp, q, xq);
      double uh = 0.0, gh[3] = {0.0, 0.0, 0.0};
      for (int i = 0; i < NN; ++i) {
        uh += nod[i] * tv.shp[i * NQ + q];
        for (int k = 0; k < 3; ++k) {
          double g = 0.0;
          for (int j = 0; j < 3; ++j) g = g + B[j][k] * tv.dshp[(i * 3 + j) * NQ + q];
          gh[k] += nod[i] * g;
        }
      }
      double u, gu[3] = {0.0, 0.0, 0.0};
      if (exact == FEMHIP_EXACT_LINEAR) {
        u = 0.0;
        for (int k = 0; k < DIM; ++k) {
          u += xq[k];
          gu[k] = 1.0;
        }
      } else {
        double s[3] = {1.0, 1.0, 1.0}, cs[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < DIM; ++k) {
          s[k] = sin(kPi * xq[k]);
          cs[k] = cos(kPi * xq[k]);
        }
        u = s[0] * s[1] * s[2];
        for (int k = 0; k < DIM; ++k) {
          double v = kPi * cs[k];
          for (int e = 0; e < DIM; ++e)
            if (e != k) v *= s[e];
          gu[k] = v;
        }
      }
      const double e = u - uh;
      double ge2 = 0.0;
      for (int k = 0; k < 3; ++k) ge2 += (gu[k] - gh[k]) * (gu[k] - gh[k]);
      vc2 += (e * e) * dv;
      gc2 += ge2 * dv;
      vmax = fmax(vmax, fabs(e));
    }
    vsum += vc2;
    gsum += gc2;
  }
  __shared__ double sv[kNormThreads], sgs[kNormThreads], sm[kNormThreads];
  sv[threadIdx.x] = vsum;
  sgs[threadIdx.x] = gsum;
  sm[threadIdx.x] = vmax;
  __syncthreads();
  for (int s = kNormThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sv[threadIdx.x] += sv[threadIdx.x + s];
      sgs[threadIdx.x] += sgs[threadIdx.x + s];
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 0] = sv[0];
    part[3 * blockIdx.x + 1] = sgs[0];
    part[3 * blockIdx.x + 2] = sm[0];
  }
}

std::vector<double> mapped_table(const RefFE &fe) {
  const int nn = fe.nnode, nq = fe.nq;
  std::vector<double> t;
  t.reserve((size_t)(nq + nn * nq * 4 + 8 * nq * 4));
  for (int q = 0; q < nq; ++q) t.push_back(fe.qw[q]);
  for (int a = 0; a < nn; ++a)
    for (int q = 0; q < nq; ++q) t.push_back(fe.shp[a][q]);
  for (int a = 0; a < nn; ++a)
    for (int j = 0; j < 3; ++j)
      for (int q = 0; q < nq; ++q) t.push_back(fe.dshp[a][j][q]);
  for (int a = 0; a < 8; ++a)
    for (int q = 0; q < nq; ++q) t.push_back(a < (1 << fe.dim) ? fe.gshp[a][q] : 0.0);
  for (int a = 0; a < 8; ++a)
    for (int j = 0; j < 3; ++j)
      for (int q = 0; q < nq; ++q) t.push_back(a < (1 << fe.dim) ? fe.gdshp[a][j][q] : 0.0);
  return t;
}

void upload_table(const femhip_space &s, DevBuf<double> &dst, hipStream_t st) {
  const std::vector<double> t = mapped_table(s.fe);
  dst.alloc(t.size());
  FH_HIP(hipMemcpyAsync(dst.p, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, st));
  FH_HIP(hipStreamSynchronize(st));
}

unsigned grid_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

bool general_integration(const femhip_operator &op) {
  if (!op.space) return false;
  if (op.space->mapped) return true;
  return op.phys.kind == FEMHIP_PHYS_POISSON &&
         (op.phys.k_kind == FEMHIP_K_QPOINT || op.phys.f_kind == FEMHIP_F_QPOINT);
}

void set_vertex_coords(femhip_space &s, const double *coords) {
  const int dim = s.mesh.dim;
  int64_t nv = 1;
  for (int k = 0; k < dim; ++k) nv *= (int64_t)s.mesh.n[k] + 1;
  const size_t count = (size_t)(nv * dim);
  for (size_t i = 0; i < count; ++i)
    FH_CHECK(std::isfinite(coords[i]), FEMHIP_ERR_INVALID, "vertex coordinates must be finite");
  s.vcoords.alloc(count);
  FH_HIP(hipMemcpy(s.vcoords.p, coords, sizeof(double) * count, hipMemcpyHostToDevice));
  s.mapped = true;
}

void quadrature_points(const femhip_space &s, double *xq_host) {
  hipStream_t st = s.ctx->stream;
  const RefFE &fe = s.fe;
  // qw[nq] | gshp[8][nq]: the leading block of the mapped table up to gshp
  std::vector<double> t;
  for (int q = 0; q < fe.nq; ++q) t.push_back(fe.qw[q]);
  for (int a = 0; a < 8; ++a)
    for (int q = 0; q < fe.nq; ++q) t.push_back(a < (1 << fe.dim) ? fe.gshp[a][q] : 0.0);
  DevBuf<double> tab, out;
  tab.alloc(t.size());
  FH_HIP(hipMemcpyAsync(tab.p, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, st));
  const size_t count = (size_t)(s.ncells * fe.nq * fe.dim);
  out.alloc(count);
  const MapMesh m = map_mesh(s);
  const double *vc = s.mapped ? s.vcoords.p : nullptr;
  const unsigned g = grid_of(s.ncells, 256);
  if (fe.dim == 2 && fe.nq == 4) hipLaunchKernelGGL((k_qpoints<2, 4>), dim3(g), dim3(256), 0, st, m, s.ncells, tab.p, vc, out.p);
  else if (fe.dim == 2 && fe.nq == 9) hipLaunchKernelGGL((k_qpoints<2, 9>), dim3(g), dim3(256), 0, st, m, s.ncells, tab.p, vc, out.p);
  else if (fe.dim == 3 && fe.nq == 8) hipLaunchKernelGGL((k_qpoints<3, 8>), dim3(g), dim3(256), 0, st, m, s.ncells, tab.p, vc, out.p);
  else if (fe.dim == 3 && fe.nq == 27) hipLaunchKernelGGL((k_qpoints<3, 27>), dim3(g), dim3(256), 0, st, m, s.ncells, tab.p, vc, out.p);
  else throw Error{FEMHIP_ERR_INVALID, "unsupported quadrature"};
  FH_HIP(hipGetLastError());
  FH_HIP(hipMemcpyAsync(xq_host, out.p, sizeof(double) * count, hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
}

void interpolate_dirichlet_mapped(const femhip_space &s, double *out_host) {
  if (s.n_fixed == 0) return;
  hipStream_t st = s.ctx->stream;
  DevBuf<double> out;
  out.alloc((size_t)s.n_fixed);
  hipLaunchKernelGGL(s.fe.dim == 2 ? k_interp_mapped<2> : k_interp_mapped<3>, dim3(grid_of(s.ncells, 256)), dim3(256),
                     0, st, face_tab(s), map_mesh(s), s.ncells, s.d_dofs.p, s.mapped ? s.vcoords.p : nullptr, out.p);
  FH_HIP(hipGetLastError());
  FH_HIP(hipMemcpyAsync(out_host, out.p, sizeof(double) * (size_t)s.n_fixed, hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
}

void upload_qpoint_values(femhip_operator &op) {
  const femhip_space &s = *op.space;
  const size_t count = (size_t)(s.ncells * s.fe.nq);
  if (op.phys.k_kind == FEMHIP_K_QPOINT) {
    FH_CHECK(op.phys.k_qp != nullptr, FEMHIP_ERR_INVALID, "k_qp is NULL for FEMHIP_K_QPOINT");
    op.k_qp.alloc(count);
    FH_HIP(hipMemcpy(op.k_qp.p, op.phys.k_qp, sizeof(double) * count, hipMemcpyHostToDevice));
  }
  if (op.phys.f_kind == FEMHIP_F_QPOINT) {
    FH_CHECK(op.phys.f_qp != nullptr, FEMHIP_ERR_INVALID, "f_qp is NULL for FEMHIP_F_QPOINT");
    op.f_qp.alloc(count);
    FH_HIP(hipMemcpy(op.f_qp.p, op.phys.f_qp, sizeof(double) * count, hipMemcpyHostToDevice));
  }
  op.phys.k_qp = op.phys.f_qp = nullptr;   // host pointers are borrowed for the call only
}

void element_matrices_mapped(femhip_operator &op) {
  femhip_space &s = *op.space;
  const RefFE &fe = s.fe;
  hipStream_t st = op.ctx->stream;
  if (op.mtab.n == 0) upload_table(s, op.mtab, st);
  const int nd = fe.nd;
  if (op.elmat.n != (size_t)(s.ncells * nd * nd)) {
    op.elmat.alloc((size_t)(s.ncells * nd * nd));
    op.elvec.alloc((size_t)(s.ncells * nd));
  }
  MapPhys ph{};
  ph.kind = op.phys.kind;
  ph.k_kind = op.phys.k_kind;
  ph.f_kind = op.phys.f_kind;
  ph.k0 = op.phys.k0;
  ph.f0 = op.phys.f0;
  if (op.phys.kind == FEMHIP_PHYS_ELASTICITY) {
    const double E = op.phys.E, nu = op.phys.nu;
    ph.lambda = (nu * E) / ((1 + nu) * (1 - 2 * nu));
    ph.mu = E / (2 * (1 + nu));
    for (int k = 0; k < 3; ++k) ph.body[k] = op.phys.body[k];
  }
  const MapMesh m = map_mesh(s);
  const double *vc = s.mapped ? s.vcoords.p : nullptr;
  const dim3 G((unsigned)s.ncells), B(64);
  const double *kc = op.k_cell.p, *kq = op.k_qp.p, *fq = op.f_qp.p;
  if (op.tab_kind == 0) {
    if (fe.dim == 2 && fe.nnode == 4)
      hipLaunchKernelGGL((k_elmat_mapped<2, 4, 4, 0>), G, B, 0, st, ph, m, s.ncells, op.mtab.p, vc, kc, kq, fq, op.elmat.p, op.elvec.p);
    else if (fe.dim == 2 && fe.nnode == 9)
      hipLaunchKernelGGL((k_elmat_mapped<2, 9, 9, 0>), G, B, 0, st, ph, m, s.ncells, op.mtab.p, vc, kc, kq, fq, op.elmat.p, op.elvec.p);
    else if (fe.dim == 3 && fe.nnode == 8)
      hipLaunchKernelGGL((k_elmat_mapped<3, 8, 8, 0>), G, B, 0, st, ph, m, s.ncells, op.mtab.p, vc, kc, kq, fq, op.elmat.p, op.elvec.p);
    else if (fe.dim == 3 && fe.nnode == 27)
      hipLaunchKernelGGL((k_elmat_mapped<3, 27, 27, 0>), G, B, 0, st, ph, m, s.ncells, op.mtab.p, vc, kc, kq, fq, op.elmat.p, op.elvec.p);
    else
      throw Error{FEMHIP_ERR_INVALID, "unsupported Poisson element"};
  } else {
    if (fe.dim == 3 && fe.nnode == 8 && fe.ncomp == 3)
      hipLaunchKernelGGL((k_elmat_mapped<3, 8, 8, 1>), G, B, 0, st, ph, m, s.ncells, op.mtab.p, vc, kc, kq, fq, op.elmat.p, op.elvec.p);
    else
      throw Error{FEMHIP_ERR_INVALID, "unsupported elasticity element"};
  }
  FH_HIP(hipGetLastError());
}

void error_norms_mapped(femhip_operator &op, const double *d_x, int exact, double out[4]) {
  femhip_space &s = *op.space;
  const RefFE &fe = s.fe;
  hipStream_t st = op.ctx->stream;
  DevBuf<double> tab, part;
  upload_table(s, tab, st);
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(4096, (s.ncells + kNormThreads - 1) / kNormThreads));
  part.alloc((size_t)3 * blocks);
  const MapMesh m = map_mesh(s);
  const double *vc = s.mapped ? s.vcoords.p : nullptr;
  const double *fx = op.fixed.p;
  const dim3 G(blocks), B(kNormThreads);
  if (fe.dim == 2 && fe.nd == 4)
    hipLaunchKernelGGL((k_error_norms_mapped<2, 4, 4>), G, B, 0, st, m, s.ncells, s.d_dofs.p, d_x, fx, tab.p, vc, exact, part.p);
  else if (fe.dim == 2 && fe.nd == 9)
    hipLaunchKernelGGL((k_error_norms_mapped<2, 9, 9>), G, B, 0, st, m, s.ncells, s.d_dofs.p, d_x, fx, tab.p, vc, exact, part.p);
  else if (fe.dim == 3 && fe.nd == 8)
    hipLaunchKernelGGL((k_error_norms_mapped<3, 8, 8>), G, B, 0, st, m, s.ncells, s.d_dofs.p, d_x, fx, tab.p, vc, exact, part.p);
  else if (fe.dim == 3 && fe.nd == 27)
    hipLaunchKernelGGL((k_error_norms_mapped<3, 27, 27>), G, B, 0, st, m, s.ncells, s.d_dofs.p, d_x, fx, tab.p, vc, exact, part.p);
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
