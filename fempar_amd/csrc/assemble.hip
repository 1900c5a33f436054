// assemble.hip -- cell-wise Galerkin integration + deterministic owner-computes
// assembly into the CSR values and RHS on gfx950.
//
// Mirrors the integrate_galerkin plugins
//   Poisson:    Sources/Tests/Serial/test_poisson/modules/poisson_discrete_integration.f90:62-141
//               (coefficient k at the quadrature point as in the heterogeneous
//               par_test_pb_bddc_poisson plugin)
//   Elasticity: Sources/Tests/Par/par_test_pb_bddc_linear_elasticity/modules/
//               sbm_irreducible_discrete_integration.i90:64-130
// followed by impose_strong_dirichlet_bcs_conforming
// (Integration/FESpace/sbm_fe_cell_iterator.i90:1610-1637) and the sparse
// assembler's insert into a fixed CSR pattern
// (LinearAlgebra/SparseMatrix/csr_sparse_matrix.f90:1179-1211, 1369-1389).
//
// On a uniform mesh every cell is a translate of the first one, so the element
// matrix factors as  elmat_c = sum_q k_c(x_q) * M_q  with M_q = detJ w_q G_q^T G_q
// fixed per space (nq small nd x nd matrices), and elvec_c = sum_q f_c(x_q) Nw_q.
//
// Every row is assembled by its owner, which sums the element entries of the
// cells around its DOF in ascending cell index and writes each value of its
// row once: scalar Q1 (the CG workloads) with one thread per row and per-row
// slot maps, other spaces (Q2, vector Q1) with one wavefront per row.  Both
// are free of atomics and deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>

#include "femhip_internal.h"

namespace femhip {

namespace {

constexpr double kPi = 3.14159265358979323846;

struct PhysDev {
  int kind, k_kind, f_kind;
  double k0, f0;
  const double *k_cell;
};

// ---------------------------------------------- owner-computes scalar Q1 ---
// Every free DOF of a scalar Q1 space sits on one vertex, and its CSR row holds
// the free DOFs of the 3^d vertices around it.  One thread per row sums, over
// the 2^d cells around its vertex in ascending cell order (the order in which
// the reference's COO append + stable sort sums duplicates), the element
// entries coupling it to each neighbour, and writes every value of the row
// once: no colours, no search, no read-modify-write.

struct VtxGrid {
  int dim;
  int n[3];    // cells per direction (1 beyond dim)
  int nv[3];   // vertices per direction (1 beyond dim)
};

__device__ __forceinline__ void vtx_coords(const VtxGrid &g, int32_t v, int a[3]) {
  a[0] = v % g.nv[0];
  a[1] = (v / g.nv[0]) % g.nv[1];
  a[2] = v / (g.nv[0] * g.nv[1]);
}

// vertex -> signed DOF id and free row -> vertex, from the cell DOF lists
template <int DIM>
__global__ void k_gather_vertices(VtxGrid g, int64_t ncells, const int32_t *__restrict__ dofs,
                                  int32_t *__restrict__ vtx_dof, int32_t *__restrict__ row_vtx) {
  constexpr int ND = 1 << DIM;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * ND) return;
  const int64_t cell = t / ND;
  const int l = (int)(t % ND);
  const int ci = (int)(cell % g.n[0]), cj = (int)((cell / g.n[0]) % g.n[1]), ck = (int)(cell / ((int64_t)g.n[0] * g.n[1]));
  const int a = ci + (l & 1), b = cj + ((l >> 1) & 1), c = DIM > 2 ? ck + ((l >> 2) & 1) : 0;
  const int32_t v = a + g.nv[0] * (b + g.nv[1] * c);
  const int32_t id = dofs[t];
  vtx_dof[v] = id;             // every cell around v writes the same id
  if (id >= 0) row_vtx[id] = v;
}

// CSR slot of each neighbour offset code (dx+1) + 3(dy+1) + 9(dz+1) of a row;
// checks that the row holds exactly the free neighbours, in ascending order
template <int DIM>
__global__ void k_gather_slots(VtxGrid g, int32_t n, const int32_t *__restrict__ vtx_dof,
                               const int32_t *__restrict__ row_vtx, const int32_t *__restrict__ rowptr,
                               const int32_t *__restrict__ col, uint8_t *__restrict__ slot, int *__restrict__ bad) {
  constexpr int K = DIM == 2 ? 9 : 27;
  constexpr int KS = DIM == 2 ? 16 : 32;   // slot row stride (bytes)
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int a[3];
  vtx_coords(g, row_vtx[r], a);
  int32_t ids[K];
  int cnt = 0;
  for (int k = 0; k < K; ++k) {
    const int d[3] = {k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1};
    bool in = true;
    for (int e = 0; e < DIM; ++e) in = in && a[e] + d[e] >= 0 && a[e] + d[e] < g.nv[e];
    ids[k] = in ? vtx_dof[(a[0] + d[0]) + g.nv[0] * ((a[1] + d[1]) + g.nv[1] * (DIM > 2 ? a[2] + d[2] : 0))] : -1;
    cnt += ids[k] >= 0;
  }
  const int32_t b = rowptr[r], len = rowptr[r + 1] - b;
  bool ok = cnt == len;
  for (int k = 0; k < KS; ++k) {
    uint8_t s = 255;
    if (k < K && ids[k] >= 0) {
      int rank = 0;
      for (int k2 = 0; k2 < K; ++k2) rank += ids[k2] >= 0 && ids[k2] < ids[k];
      ok = ok && rank < len && col[b + rank] == ids[k];   // col read only inside the row
      s = (uint8_t)rank;
    }
    slot[(int64_t)r * KS + k] = s;
  }
  if (!ok) atomicOr(bad, 1);
}

// Every present neighbour `code` of every row must sit at column row + off[code].
struct CodeOffsets {
  int32_t off[27];
};
template <int DIM>
__global__ void k_check_code_offsets(int32_t n, const uint8_t *__restrict__ slot, const int32_t *__restrict__ rowptr,
                                     const int32_t *__restrict__ col, CodeOffsets co, int *__restrict__ bad) {
  constexpr int K = DIM == 2 ? 9 : 27;
  constexpr int KS = DIM == 2 ? 16 : 32;
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t b = rowptr[r];
  bool ok = true;
  for (int k = 0; k < K; ++k) {
    const uint8_t s = slot[(int64_t)r * KS + k];
    if (s != 255) ok = ok && col[b + s] - r == co.off[k];
  }
  if (!ok) atomicOr(bad, 1);
}

// On a uniform mesh the d-th coordinate of quadrature point q of cell (i,j,k)
// depends only on (i_d, the d-th 1-D Gauss index of q), so the trigonometric
// factors of the coefficient and source are tabulated per direction:
// trig[off[d] + i_d * nq1 + a] = {sin(pi x), cos(pi x), sin(2 pi x), cos(2 pi x)}.
struct TrigTab {
  const double4 *t;
  int off[3];
};

// k and f at quadrature point q (NQ1 Gauss points per direction, first index
// fastest) from the tables: k = 1 + 0.5 sin 2pi x sin 2pi y sin 2pi z (or the
// cell's value, or constant), f = -div(k grad u) for u = sin pi x sin pi y
// sin pi z (the coefficient variant of poisson_discrete_integration.f90)
template <int DIM, int NQ1 = 2>
__device__ __forceinline__ void coef_tab(const PhysDev &ph, const TrigTab &tt, const int ijk[3], int q, int64_t cell,
                                         double &k, double &f) {
  double su[3] = {1, 1, 1}, cu[3] = {0, 0, 0}, s2[3] = {1, 1, 1}, c2[3] = {0, 0, 0};
  int qr = q;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int qd = qr % NQ1;
    qr /= NQ1;
    const double4 v = tt.t[tt.off[d] + ijk[d] * NQ1 + qd];
    su[d] = v.x;
    cu[d] = v.y;
    s2[d] = v.z;
    c2[d] = v.w;
  }
  if (ph.k_kind == FEMHIP_K_SIN3) k = 1.0 + 0.5 * (s2[0] * s2[1] * s2[2]);
  else if (ph.k_kind == FEMHIP_K_CELL) k = ph.k_cell[cell];
  else k = ph.k0;
  if (ph.f_kind == FEMHIP_F_MANUFACTURED) {
    f = k * DIM * kPi * kPi * (su[0] * su[1] * su[2]);
    if (ph.k_kind == FEMHIP_K_SIN3) {
      double dot = kPi * c2[0] * s2[1] * s2[2] * kPi * cu[0] * su[1] * su[2] +
                   kPi * s2[0] * c2[1] * s2[2] * kPi * su[0] * cu[1] * su[2];
      if (DIM == 3) dot += kPi * s2[0] * s2[1] * c2[2] * kPi * su[0] * su[1] * cu[2];
      f -= dot;
    }
  } else if (ph.f_kind == FEMHIP_F_CELL_SIN) {
    f = ph.f0 * k * DIM * kPi * kPi * (su[0] * su[1] * su[2]);
  } else {
    f = ph.f0;
  }
}

// ----------------------------------------- owner-computes, any Lagrangian ---
// One wavefront per free row (Q2 scalar, vector Q1, and scalar Q1 where the
// vertex maps do not apply).  The row's DOF sits on the vef of its
// representative occurrence (cell, local DOF); the cells around that vef are
// visited in ascending cell index (the order in which the reference's COO
// append + stable sort sums duplicates, base_sparse_matrix.f90:3469-3513,
// 4121-4829), and each adds its element row for the row's local DOF: lane j
// forms entry (l, j) and adds it to the row's value at the CSR slot of column
// dofs[cell, j] (found by binary search in the row's sorted columns, kept in
// LDS with the values); a fixed column lifts e g_j into the RHS in column
// order, as impose_strong_dirichlet_bcs does (sbm_fe_cell_iterator.i90:
// 1610-1637).  The row is written once, coalesced: no clearing pass, no
// colours, no read-modify-write in HBM.
// MODE 0: Poisson, entry = sum_q k(x_q) M_q[l][j], RHS sum_q f(x_q) Nw_q[l],
// k and f from the per-direction sin/cos tables at the NQ1 1-D Gauss points;
// MODE 1: one element matrix and vector for every cell (elasticity);
// MODE 2: each cell's own element matrix and vector, formed beforehand
// (tab = cells x ND x ND row-major, evec = cells x ND; cells with their own map
// or coefficients given at the quadrature points).
constexpr int kOwnMaxRow = 128;
template <int DIM, int ND, int NQ1, int MODE>
__global__ __launch_bounds__(64) void k_assemble_owner(PhysDev ph, TrigTab tt, FaceTab t, int32_t n,
                                                       const int64_t *__restrict__ rep,
                                                       const int32_t *__restrict__ dofs,
                                                       const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ col, const double *__restrict__ tab,
                                                       const double *__restrict__ fixed, double *__restrict__ val,
                                                       double *__restrict__ rhs, const double *__restrict__ evec) {
  constexpr int NQ = MODE == 0 ? (DIM == 2 ? NQ1 * NQ1 : NQ1 * NQ1 * NQ1) : 1;
  __shared__ int32_t scol[kOwnMaxRow];
  __shared__ double sacc[kOwnMaxRow];
  __shared__ double sk[NQ], sf[NQ];
  __shared__ double sfix[ND];
  const double *M = tab;                         // MODE 0: M[NQ][ND][ND] | Nw[NQ][ND]
  const double *Nw = tab + NQ * ND * ND;
  const double *E = tab, *F = tab + ND * ND;     // MODE 1: element matrix (row-major), vector
  const int lane = threadIdx.x;
  const int32_t r = blockIdx.x;
  const int32_t b0 = rowptr[r], len = rowptr[r + 1] - b0;
  for (int i = lane; i < kOwnMaxRow; i += 64) {
    scol[i] = i < len ? col[b0 + i] : INT_MAX;
    sacc[i] = 0.0;
  }
  const int64_t c0 = rep[r];
  const int64_t cell0 = c0 / ND;
  const int l0 = (int)(c0 % ND), f = t.face_of[l0], pos = t.pos_of[l0];
  int ijk[3];
  cell_ijk(t, cell0, ijk);
  int ax[3] = {0, 0, 0}, na = 0;
  for (int k = 0; k < DIM; ++k)
    if (!((t.type[f] >> k) & 1)) ax[na++] = k;
  double bacc = 0.0;
  __syncthreads();
  for (int b = 0; b < (1 << na); ++b) {   // ascending cell index
    int cj[3] = {ijk[0], ijk[1], ijk[2]}, anc = 0;
    bool ok = true;
    for (int e = 0; e < na; ++e) {
      const int k = ax[e], w = ijk[k] + ((t.anchor[f] >> k) & 1);
      cj[k] = w - 1 + ((b >> e) & 1);
      ok = ok && cj[k] >= 0 && cj[k] < t.n[k];
      anc |= (w - cj[k]) << k;
    }
    if (!ok) continue;   // the same for the whole block (one row)
    const int lc = t.lst[t.ptr[t.face_by[t.type[f]][anc]] + pos];   // the row's DOF in this cell
    const int64_t cell = cell_index(t, cj);
    if (MODE == 0) {
      if (lane < NQ) coef_tab<DIM, NQ1>(ph, tt, cj, lane, cell, sk[lane], sf[lane]);
      __syncthreads();
    }
    if (lane < ND) {
      double e;
      if (MODE == 0) {
        e = 0.0;
        for (int q = 0; q < NQ; ++q) e += sk[q] * M[(q * ND + lc) * ND + lane];
      } else if (MODE == 1) {
        e = E[lc * ND + lane];
      } else {
        e = tab[(cell * ND + lc) * ND + lane];
      }
      const int32_t gj = dofs[cell * ND + lane];
      double fx = 0.0;
      if (gj >= 0) {
        int lo = 0;   // = the number of the row's columns below gj (fixed steps, len <= 128)
#pragma unroll
        for (int h = kOwnMaxRow / 2; h > 0; h >>= 1)
          if (lo + h <= len && scol[lo + h - 1] < gj) lo += h;
        sacc[lo] += e;   // the cell's columns are distinct: one lane per slot
      } else if (fixed) {
        fx = e * fixed[-gj - 1];
      }
      sfix[lane] = fx;
    }
    __syncthreads();
    if (lane == 0) {
      double be;
      if (MODE == 0) {
        be = 0.0;
        for (int q = 0; q < NQ; ++q) be += sf[q] * Nw[q * ND + lc];
      } else if (MODE == 1) {
        be = F[lc];
      } else {
        be = evec[cell * ND + lc];
      }
      for (int j = 0; j < ND; ++j) be -= sfix[j];
      bacc += be;
    }
    __syncthreads();
  }
  for (int i = lane; i < len; i += 64) val[b0 + i] = sacc[i];
  if (lane == 0) rhs[r] = bacc;
}

// ----------------------------------------------- 3-D Q2: element matrices ---
// Phase 1 forms every cell's element matrix as one FP64 GEMM on the matrix
// cores: elmat[cell][l j] = sum_q k_cell(q) M_q[l][j], i.e. (cells x 28) k
// times (28 x 736) M (27 points and 729 entries, zero-padded), in
// v_mfma_f64_16x16x4f64 tiles of 16 cells x 16 entries; a wave holds the k of
// its 16 cells (7 K-steps) in registers and sweeps the 46 column tiles.  The
// element vectors elvec[cell][l] = sum_q f_cell(q) Nw_q[l] are summed in q
// order on the vector ALUs.  Operand layouts of the instruction: A[m][k] in
// lane m + 16 k, B[k][n] in lane n + 16 k, D[m][n] in lane n + 16 (m % 4),
// slot m / 4.
constexpr int kQ2Nd = 27, kQ2Nq = 27, kQ2Ent = kQ2Nd * kQ2Nd, kQ2Tiles = (kQ2Ent + 15) / 16, kQ2Steps = 7;
__global__ __launch_bounds__(64) void k_elmat_q2(PhysDev ph, TrigTab tt, FaceTab t, int64_t ncells,
                                                 const double *__restrict__ tab, double *__restrict__ elmat,
                                                 double *__restrict__ elvec) {
  __shared__ double sf[16][kQ2Nq];
  const double *M = tab;                          // M[q][l * 27 + j]
  const double *Nw = tab + kQ2Nq * kQ2Ent;        // Nw[q][l]
  const int lane = threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.x * 16;
  const int m = lane & 15, kq = lane >> 4;
  double a[kQ2Steps];
  {
    const int64_t cell = c0 + m;
    int ijk[3] = {0, 0, 0};
    if (cell < ncells) cell_ijk(t, cell, ijk);
#pragma unroll
    for (int st = 0; st < kQ2Steps; ++st) {
      const int q = 4 * st + kq;
      double kv = 0.0, fv = 0.0;
      if (cell < ncells && q < kQ2Nq) coef_tab<3, 3>(ph, tt, ijk, q, cell, kv, fv);
      a[st] = kv;
    }
  }
  typedef double v4d __attribute__((ext_vector_type(4)));
  for (int tile = 0; tile < kQ2Tiles; ++tile) {
    const int colb = tile * 16 + m;             // this lane's B column
    v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int st = 0; st < kQ2Steps; ++st) {
      const int q = 4 * st + kq;
      const double b = (q < kQ2Nq && colb < kQ2Ent) ? M[q * kQ2Ent + colb] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[st], b, acc, 0, 0, 0);
    }
    const int col = tile * 16 + (lane & 15);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t cell = c0 + 4 * i + (lane >> 4);
      if (cell < ncells && col < kQ2Ent) elmat[cell * kQ2Ent + col] = acc[i];
    }
  }
  // element vectors: f at the points of the 16 cells, then sum_q f Nw in q order
  for (int p = lane; p < 16 * kQ2Nq; p += 64) {
    const int cm = p / kQ2Nq, q = p % kQ2Nq;
    const int64_t cell = c0 + cm;
    double kv = 0.0, fv = 0.0;
    if (cell < ncells) {
      int ijk[3];
      cell_ijk(t, cell, ijk);
      coef_tab<3, 3>(ph, tt, ijk, q, cell, kv, fv);
    }
    sf[cm][q] = fv;
  }
  __syncthreads();
  for (int p = lane; p < 16 * kQ2Nd; p += 64) {
    const int cm = p / kQ2Nd, l = p % kQ2Nd;
    const int64_t cell = c0 + cm;
    if (cell >= ncells) continue;
    double be = 0.0;
    for (int q = 0; q < kQ2Nq; ++q) be += sf[cm][q] * Nw[q * kQ2Nd + l];
    elvec[cell * kQ2Nd + l] = be;
  }
}

// Phase 2, one wavefront per free row: the cells around the row's vef in
// ascending index (at most 8) and the row's local DOF in each; every (cell, j)
// pair at once over the lanes: entry elmat[cell][l j], its column's CSR slot
// (binary search in the row's sorted columns in LDS), or its Dirichlet lift
// e g_j; then each slot sums its contributions in ascending cell order (the
// reference's duplicate-summation order) and the row is written once; the RHS
// is sum over cells of (elvec - the lifts in column order).
// Each wavefront works on its own row and LDS region; it orders its lanes'
// LDS traffic with wavefront barriers only, so the block's waves never wait
// for each other.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// rows (wavefronts) per block: the 64^3 Q2 assembly took 6.1-6.3 ms with 1,
// 7.2 ms with 4 (block barriers or wavefront barriers alike)
constexpr int kQ2Rows = 1;
__global__ __launch_bounds__(64 * kQ2Rows) void k_gather_q2(FaceTab t, int32_t n, const int64_t *__restrict__ rep,
                                                            const int32_t *__restrict__ dofs,
                                                            const int32_t *__restrict__ rowptr,
                                                            const int32_t *__restrict__ col,
                                                            const double *__restrict__ elmat,
                                                            const double *__restrict__ elvec,
                                                            const double *__restrict__ fixed, double *__restrict__ val,
                                                            double *__restrict__ rhs) {
  constexpr int kCells = 8, kPairs = kCells * kQ2Nd;
  __shared__ int32_t scol_[kQ2Rows][kOwnMaxRow];
  __shared__ int16_t slist_[kQ2Rows][kOwnMaxRow][kCells];   // pair feeding (slot, cell), -1 if none
  __shared__ double contrib_[kQ2Rows][kPairs];
  __shared__ double lift_[kQ2Rows][kPairs];
  __shared__ int64_t scell_[kQ2Rows][kCells];
  __shared__ int slc_[kQ2Rows][kCells];
  __shared__ double sbe_[kQ2Rows][kCells];
  __shared__ int snc_[kQ2Rows];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int32_t *scol = scol_[w];
  int16_t(*slist)[kCells] = slist_[w];
  double *contrib = contrib_[w], *lift = lift_[w], *sbe = sbe_[w];
  int64_t *scell = scell_[w];
  int *slc = slc_[w];
  const int32_t r = blockIdx.x * kQ2Rows + w;
  if (r >= n) return;   // no block-wide barrier below
  const int32_t b0 = rowptr[r], len = rowptr[r + 1] - b0;
  for (int i = lane; i < len; i += 64) {
    scol[i] = col[b0 + i];
    for (int c = 0; c < kCells; ++c) slist[i][c] = -1;
  }
  {   // the cells around the row's vef, ascending (lane b tries cell b), and the row's DOF in each
    const int64_t c0 = rep[r];
    const int l0 = (int)(c0 % kQ2Nd), f = t.face_of[l0], pos = t.pos_of[l0];
    int ijk[3];
    cell_ijk(t, c0 / kQ2Nd, ijk);
    // axis k not tangent to the vef is its e-th such axis, e = the number of
    // such axes below k; lane b takes bit e of b along it (all indices static,
    // so nothing goes to scratch: 2.89 -> 2.52 ms at 64^3 with the load below)
    const int nt = ~t.type[f] & 7;
    const int b = lane;
    int cj[3], anc = 0;
    bool ok = b < (1 << __popc(nt));
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      cj[k] = ijk[k];
      if ((nt >> k) & 1) {
        const int e = __popc(nt & ((1 << k) - 1));
        const int wk = ijk[k] + ((t.anchor[f] >> k) & 1);
        cj[k] = wk - 1 + ((b >> e) & 1);
        ok = ok && cj[k] >= 0 && cj[k] < t.n[k];
        anc |= (wk - cj[k]) << k;
      }
    }
    const uint64_t mask = __ballot(ok);
    if (ok) {
      const int at = __popcll(mask & ((1ull << lane) - 1));   // ascending b = ascending cell index
      scell[at] = cell_index(t, cj);
      slc[at] = t.lst[t.ptr[t.face_by[t.type[f]][anc]] + pos];
    }
    if (lane == 0) snc_[w] = __popcll(mask);
  }
  wave_sync();
  const int nc = snc_[w];
  // the element-vector entries go out with the element rows, not after the sums
  const double ebv = lane < nc ? elvec[scell[lane] * kQ2Nd + slc[lane]] : 0.0;
  constexpr int kRounds = (kPairs + 63) / 64;
  double ev[kRounds];
  int32_t gv[kRounds];
#pragma unroll
  for (int k = 0; k < kRounds; ++k) {   // every load of the lane in flight before the searches
    const int p = lane + 64 * k;
    ev[k] = 0.0;
    gv[k] = -1;
    if (p < nc * kQ2Nd) {
      const int ci = p / kQ2Nd, j = p % kQ2Nd;
      const int64_t cell = scell[ci];
      ev[k] = elmat[cell * kQ2Ent + slc[ci] * kQ2Nd + j];
      gv[k] = dofs[cell * kQ2Nd + j];
    }
  }
#pragma unroll
  for (int k = 0; k < kRounds; ++k) {
    const int p = lane + 64 * k;
    if (p >= nc * kQ2Nd) continue;
    const int ci = p / kQ2Nd;
    const double e = ev[k];
    const int32_t gj = gv[k];
    double lf = 0.0;
    if (gj >= 0) {
      // first slot whose column is not below gj, in a fixed 7 steps (len <=
      // 128), so the lane's independent searches overlap
      int lo = 0;   // = the number of the row's columns below gj
#pragma unroll
      for (int h = kOwnMaxRow / 2; h > 0; h >>= 1)
        if (lo + h <= len && scol[lo + h - 1] < gj) lo += h;
      slist[lo][ci] = (int16_t)p;   // a cell's columns are distinct: one pair per (slot, cell)
      contrib[p] = e;
    } else if (fixed) {
      lf = e * fixed[-gj - 1];
    }
    lift[p] = lf;
  }
  wave_sync();
  for (int s = lane; s < len; s += 64) {
    double v = 0.0;
    for (int c = 0; c < nc; ++c) {
      const int p = slist[s][c];
      if (p >= 0) v += contrib[p];
    }
    val[b0 + s] = v;
  }
  if (lane < nc) {
    double be = ebv;
    for (int j = 0; j < kQ2Nd; ++j) be -= lift[lane * kQ2Nd + j];
    sbe[lane] = be;
  }
  wave_sync();
  if (lane == 0) {
    double b = 0.0;
    for (int c = 0; c < nc; ++c) b += sbe[c];
    rhs[r] = b;
  }
}

// The solver's symmetric-diagonal copy, written by the assembly as well when
// v64 is set: V[kidx[code]][row] = the row's value for neighbour `code` (0 for
// a fixed neighbour), for the codes of the stored (upper) diagonals.
struct SymOut {
  double *v64;
  int64_t stride;
  int32_t kp;      // stored diagonals
  int8_t kidx[32];
};

// Outputs of one row from its neighbour values acc[code] and source sum bacc:
// the CSR values staged in sv from `base` (the row's slots), fixed neighbours
// lifted into the RHS, the RHS entry, and the solver's symmetric copy.
template <int DIM>
__device__ __forceinline__ void finish_row(const VtxGrid &g, int32_t r, const int (&a)[3],
                                           const double (&acc)[DIM == 2 ? 9 : 27], double bacc,
                                           const int32_t *__restrict__ vtx_dof, const uint8_t *__restrict__ slot,
                                           int32_t base, double *sv, const double *__restrict__ fixed,
                                           double *__restrict__ rhs, const SymOut &so) {
  constexpr int K = DIM == 2 ? 9 : 27, KS = DIM == 2 ? 16 : 32;
  const uint8_t *sr = slot + (int64_t)r * KS;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint8_t s = sr[k];
    if (s != 255) {
      sv[base + s] = acc[k];
    } else if (fixed) {   // a fixed neighbour lifts its value into the RHS
      const int d[3] = {k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1};
      bool in = true;
      for (int e = 0; e < DIM; ++e) in = in && a[e] + d[e] >= 0 && a[e] + d[e] < g.nv[e];
      if (in) {
        const int32_t id = vtx_dof[(a[0] + d[0]) + g.nv[0] * ((a[1] + d[1]) + g.nv[1] * (DIM > 2 ? a[2] + d[2] : 0))];
        if (id < 0) bacc -= acc[k] * fixed[-id - 1];
      }
    }
  }
  rhs[r] = bacc;
  if (so.v64) {   // stores of one diagonal are coalesced over the block's rows
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int d = so.kidx[k];
      if (d >= 0) so.v64[sym_at(d, r, so.stride, so.kp)] = sr[k] != 255 ? acc[k] : 0.0;
    }
  }
}

// One thread per free row; a block owns RB consecutive rows, whose CSR values
// are contiguous: they are staged in LDS and stored with coalesced writes.
template <int DIM, int RB>
__global__ __launch_bounds__(RB) void k_assemble_gather(
    VtxGrid g, PhysDev ph, TrigTab tt, int32_t n, const int32_t *__restrict__ row_vtx,
    const int32_t *__restrict__ vtx_dof, const uint8_t *__restrict__ slot, const int32_t *__restrict__ rowptr,
    const double *__restrict__ tab, const double *__restrict__ fixed, double *__restrict__ val,
    double *__restrict__ rhs, SymOut so) {
  constexpr int ND = 1 << DIM, NQ = ND, K = DIM == 2 ? 9 : 27;
  __shared__ double sv[RB * K];
  const double *M = tab;                 // M[NQ][ND][ND]
  const double *Nw = tab + NQ * ND * ND; // Nw[NQ][ND]
  const int32_t r0 = blockIdx.x * RB, r = r0 + threadIdx.x;
  const int32_t b0 = rowptr[r0], b1 = rowptr[min(n, r0 + RB)];
  if (r < n) {
  int a[3];
  vtx_coords(g, row_vtx[r], a);
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  double bacc = 0.0;
#pragma unroll
  for (int e = 0; e < ND; ++e) {          // cells around the vertex, ascending index
    const int ex = e & 1, ey = (e >> 1) & 1, ez = (e >> 2) & 1;
    const int ci = a[0] - 1 + ex, cj = a[1] - 1 + ey, ck = DIM > 2 ? a[2] - 1 + ez : 0;
    if (ci < 0 || ci >= g.n[0] || cj < 0 || cj >= g.n[1] || (DIM > 2 && (ck < 0 || ck >= g.n[2]))) continue;
    const int64_t cell = ci + (int64_t)g.n[0] * (cj + (int64_t)g.n[1] * ck);
    const int ijk[3] = {ci, cj, ck};
    const int il = (1 - ex) + 2 * (1 - ey) + (DIM > 2 ? 4 * (1 - ez) : 0);   // the vertex in this cell
    double kv[NQ], fv[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) coef_tab<DIM>(ph, tt, ijk, q, cell, kv[q], fv[q]);
    double be = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) be += fv[q] * Nw[q * ND + il];
    bacc += be;
#pragma unroll
    for (int j = 0; j < ND; ++j) {
      const int dx = ex + (j & 1) - 1, dy = ey + ((j >> 1) & 1) - 1, dz = DIM > 2 ? ez + ((j >> 2) & 1) - 1 : -1;
      const int code = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) v += kv[q] * M[(q * ND + il) * ND + j];
      acc[code] += v;
    }
  }
  finish_row<DIM>(g, r, a, acc, bacc, vtx_dof, slot, rowptr[r] - b0, sv, fixed, rhs, so);
  }
  __syncthreads();
  for (int32_t e = threadIdx.x; e < b1 - b0; e += RB) val[b0 + e] = sv[e];   // rows <= K entries each
}

// 3-D scalar Q1 with the row values formed on the matrix cores.  For a tile of
// 16 rows, acc[row][code] = sum over (cell e around the vertex, point q) of
// k_e(q) W[8 e + q][code] and the source sum b[row] = sum f_e(q) Nw[q][7 - e]
// are two FP64 MFMA chains (v_mfma_f64_16x16x4: A = k or f of (row, e, q),
// B = W or the Nw column), with W[8 e + q][code] = M[q][7 - e][j] for the
// neighbour code of local vertex j seen from cell e and 0 otherwise.  K runs
// over (e, q) with e ascending (ascending cell index) and every 4-wide step
// inside one cell, so entries (r, c) and (c, r) see the same products in the
// same order: the matrix stays bitwise symmetric.  A block holds 128 rows in
// 2 waves of 4 tiles; the results pass through LDS to one thread per row.
// Operand layouts of the instruction (checked on the device): A[m][k] in lane
// m + 16 k, B[k][n] in lane n + 16 k, D[m][n] in lane n + 16 (m % 4), slot m / 4.
constexpr int kMfmaRows = 128;
typedef double v4d __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kMfmaRows) void k_assemble_gather_mfma(
    VtxGrid g, PhysDev ph, TrigTab tt, int32_t n, const int32_t *__restrict__ row_vtx,
    const int32_t *__restrict__ vtx_dof, const uint8_t *__restrict__ slot, const int32_t *__restrict__ rowptr,
    const double *__restrict__ tab, const double *__restrict__ fixed, double *__restrict__ val,
    double *__restrict__ rhs, SymOut so) {
  constexpr int ND = 8, NQ = 8, K = 27, KK = 64, WC = 32;
  __shared__ double lds[kMfmaRows * K];   // W (64 x 32), then the row values, then the staged CSR values
  __shared__ double sb[kMfmaRows];
  const double *M = tab;                  // M[NQ][ND][ND]
  const double *Nw = tab + NQ * ND * ND;  // Nw[NQ][ND]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int32_t r0 = blockIdx.x * kMfmaRows;
  for (int i = tid; i < KK * WC; i += kMfmaRows) {   // W; column 31 carries Nw for the source chain
    const int kk = i / WC, code = i % WC, e = kk >> 3, q = kk & 7;
    double w = 0.0;
    if (code < K) {
      const int jx = code % 3 - (e & 1), jy = (code / 3) % 3 - ((e >> 1) & 1), jz = code / 9 - (e >> 2);
      if (jx >= 0 && jx <= 1 && jy >= 0 && jy <= 1 && jz >= 0 && jz <= 1)
        w = M[(q * ND + (7 - e)) * ND + jx + 2 * jy + 4 * jz];
    } else if (code == WC - 1) {
      w = Nw[q * ND + (7 - e)];
    }
    lds[i] = w;
  }
  int ax[4][3];
  bool ok[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {   // the lane's row in each of the wave's four tiles
    const int32_t row = r0 + wv * 64 + mt * 16 + (lane & 15);
    ok[mt] = row < n;
    ax[mt][0] = ax[mt][1] = ax[mt][2] = 0;
    if (ok[mt]) vtx_coords(g, row_vtx[row], ax[mt]);
  }
  __syncthreads();
  v4d acc[4][2], accf[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    accf[mt] = v4d{0.0, 0.0, 0.0, 0.0};
    acc[mt][0] = acc[mt][1] = accf[mt];
  }
  for (int ks = 0; ks < KK / 4; ++ks) {
    const int kk = ks * 4 + (lane >> 4), e = kk >> 3, q = kk & 7;
    const double w0 = lds[kk * WC + (lane & 15)], w1 = lds[kk * WC + 16 + (lane & 15)];
    const double wf = (lane & 15) == 0 ? lds[kk * WC + WC - 1] : 0.0;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      double kq = 0.0, fq = 0.0;
      const int ci = ax[mt][0] - 1 + (e & 1), cj = ax[mt][1] - 1 + ((e >> 1) & 1), ck = ax[mt][2] - 1 + (e >> 2);
      if (ok[mt] && ci >= 0 && ci < g.n[0] && cj >= 0 && cj < g.n[1] && ck >= 0 && ck < g.n[2]) {
        const int64_t cell = ci + (int64_t)g.n[0] * (cj + (int64_t)g.n[1] * ck);
        const int ijk[3] = {ci, cj, ck};
        coef_tab<3>(ph, tt, ijk, q, cell, kq, fq);
      }
      acc[mt][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(kq, w0, acc[mt][0], 0, 0, 0);
      acc[mt][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(kq, w1, acc[mt][1], 0, 0, 0);
      accf[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(fq, wf, accf[mt], 0, 0, 0);
    }
  }
  __syncthreads();   // W is no longer read: lds takes the row values
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int lr = wv * 64 + mt * 16 + (lane >> 4) + 4 * i;
      const int c0 = lane & 15;
      lds[lr * K + c0] = acc[mt][0][i];
      if (16 + c0 < K) lds[lr * K + 16 + c0] = acc[mt][1][i];
      if (c0 == 0) sb[lr] = accf[mt][i];
    }
  __syncthreads();
  const int32_t r = r0 + tid;
  double av[K];
#pragma unroll
  for (int k = 0; k < K; ++k) av[k] = lds[tid * K + k];
  const double bacc = sb[tid];
  int a[3] = {0, 0, 0};
  if (r < n) vtx_coords(g, row_vtx[r], a);
  const int32_t b0 = rowptr[r0], b1 = rowptr[min(n, r0 + kMfmaRows)];
  __syncthreads();   // lds now stages the block's CSR values
  if (r < n) finish_row<3>(g, r, a, av, bacc, vtx_dof, slot, rowptr[r] - b0, lds, fixed, rhs, so);
  __syncthreads();
  for (int32_t e = tid; e < b1 - b0; e += kMfmaRows) val[b0 + e] = lds[e];   // rows <= K entries each
}

// k alone at quadrature point q (the FP32 integration needs no source)
template <int DIM, int NQ1 = 2>
__device__ __forceinline__ double coef_k(const PhysDev &ph, const TrigTab &tt, const int ijk[3], int q,
                                         int64_t cell) {
  if (ph.k_kind == FEMHIP_K_CELL) return ph.k_cell[cell];
  if (ph.k_kind != FEMHIP_K_SIN3) return ph.k0;
  double s[3] = {1, 1, 1};
  int qr = q;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int qd = qr % NQ1;
    qr /= NQ1;
    s[d] = tt.t[tt.off[d] + ijk[d] * NQ1 + qd].z;
  }
  return 1.0 + 0.5 * (s[0] * s[1] * s[2]);
}

// Stored (upper) diagonals of the FP32 symmetric-diagonal copy: diagonal d
// holds neighbour code code[d] (-1: none) at column offset off[d].
struct Sym32Out {
  float *v32;
  int64_t stride;
  int32_t nd;
  int8_t code[16];
  int32_t off[16];
};

// FP32 integration of the refined solver's matrix (scalar Q1, symmetric-
// diagonal layout) on the matrix cores with FP32 accumulation.  For a tile of
// 16 rows, acc[row][d] = sum over the (cell e around the row's vertex, point q)
// pairs of k_e(q) W[NQ e + q][d], with W[NQ e + q][d] = M_q[row's local vertex
// in e][local vertex of the neighbour on diagonal d] (0 when that neighbour is
// not a vertex of e): one v_mfma_f32_16x16x4f32 per 4 pairs and tile, the
// MFMA's FP32 accumulator holding the sum.  Each stored value is then scaled
// by S_r S_c (S = diag(A)^-1/2 of the FP64 matrix; none without Jacobi) and
// written once; a fixed or absent neighbour and the padding rows get 0.
// Operand layouts of the instruction: A[m][k] in lane m + 16 k, B[k][n] in
// lane n + 16 k, D[m][n] in lane n + 16 (m / 4), slot m % 4.
typedef float v4f __attribute__((ext_vector_type(4)));
template <int DIM>
__global__ __launch_bounds__(kMfmaRows) void k_assemble_sym32_mfma(
    VtxGrid g, PhysDev ph, TrigTab tt, int32_t n, int64_t rows, const int32_t *__restrict__ row_vtx,
    const uint8_t *__restrict__ slot, const double *__restrict__ tab, const double *__restrict__ sc, Sym32Out so) {
  constexpr int ND = 1 << DIM, NQ = ND, KK = ND * NQ, KS = DIM == 2 ? 16 : 32;
  __shared__ float w[KK * 16];
  __shared__ float out[kMfmaRows * 17];   // row values, padded against bank conflicts
  const double *M = tab;                  // M[NQ][ND][ND]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t r0 = (int64_t)blockIdx.x * kMfmaRows;
  for (int i = tid; i < KK * 16; i += kMfmaRows) {
    const int kk = i / 16, d = i % 16, e = kk / NQ, q = kk % NQ;
    float v = 0.f;
    const int code = d < so.nd ? so.code[d] : -1;
    if (code >= 0) {
      const int jx = code % 3 - (e & 1), jy = (code / 3) % 3 - ((e >> 1) & 1);
      const int jz = DIM > 2 ? code / 9 - (e >> 2) : 0;
      if (jx >= 0 && jx <= 1 && jy >= 0 && jy <= 1 && jz >= 0 && jz <= 1)
        v = (float)M[(q * ND + (ND - 1 - e)) * ND + jx + 2 * jy + 4 * jz];
    }
    w[i] = v;
  }
  int ax[4][3];
  bool ok[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {   // the lane's row in each of the wave's four tiles
    const int64_t row = r0 + wv * 64 + mt * 16 + (lane & 15);
    ok[mt] = row < n;
    ax[mt][0] = ax[mt][1] = ax[mt][2] = 0;
    if (ok[mt]) vtx_coords(g, row_vtx[row], ax[mt]);
  }
  __syncthreads();
  v4f acc[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) acc[mt] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int ks = 0; ks < KK / 4; ++ks) {
    const int kk = ks * 4 + (lane >> 4), e = kk / NQ, q = kk % NQ;
    const float wb = w[kk * 16 + (lane & 15)];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      float kq = 0.f;
      const int ci = ax[mt][0] - 1 + (e & 1), cj = ax[mt][1] - 1 + ((e >> 1) & 1);
      const int ck = DIM > 2 ? ax[mt][2] - 1 + (e >> 2) : 0;
      if (ok[mt] && ci >= 0 && ci < g.n[0] && cj >= 0 && cj < g.n[1] && (DIM == 2 || (ck >= 0 && ck < g.n[2]))) {
        const int64_t cell = ci + (int64_t)g.n[0] * (cj + (int64_t)g.n[1] * ck);
        const int ijk[3] = {ci, cj, ck};
        kq = (float)coef_k<DIM>(ph, tt, ijk, q, cell);
      }
      acc[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(kq, wb, acc[mt], 0, 0, 0);
    }
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(wv * 64 + mt * 16 + 4 * (lane >> 4) + i) * 17 + (lane & 15)] = acc[mt][i];
  __syncthreads();
  const int64_t r = r0 + tid;
  if (r >= rows) return;
  const double sr = (sc && r < n) ? sc[r] : 1.0;
  for (int d = 0; d < so.nd; ++d) {
    float a = 0.f;
    const int code = so.code[d];
    if (r < n && code >= 0 && slot[r * KS + code] != 255) {
      a = out[tid * 17 + d];
      if (sc) a = (float)((double)a * (sr * sc[min(r + (int64_t)so.off[d], (int64_t)n - 1)]));
    }
    so.v32[sym_at(d, r, so.stride, so.nd)] = a;
  }
}

__global__ void k_spmv_csr(int32_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                           const double *__restrict__ val, const double *__restrict__ x, double *__restrict__ y) {
  int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double acc = 0.0;
  for (int32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) acc += val[k] * x[col[k]];
  y[r] = acc;
}


void build_trig(femhip_operator &op);

void launch_q2(femhip_operator &op, const PhysDev &ph, const double *d_fixed) {
  femhip_space &s = *op.space;
  if (op.n <= 0) return;
  FH_CHECK(s.fe.order == 2 && s.fe.ncomp == 1 && s.fe.dim == 3, FEMHIP_ERR_INVALID, "not a scalar 3-D Q2 space");
  if (op.trig.n == 0) build_trig(op);
  const int64_t nc = s.ncells;   // rows of 5^3 = 125 <= kOwnMaxRow columns
  if (op.elmat.n != (size_t)(nc * kQ2Ent)) {
    op.elmat.alloc((size_t)(nc * kQ2Ent));
    op.elvec.alloc((size_t)(nc * kQ2Nd));
  }
  const TrigTab tt{reinterpret_cast<const double4 *>(op.trig.p), {op.trig_off[0], op.trig_off[1], op.trig_off[2]}};
  const FaceTab t = face_tab(s);
  hipLaunchKernelGGL(k_elmat_q2, dim3((unsigned)((nc + 15) / 16)), dim3(64), 0, op.ctx->stream, ph, tt, t, nc, op.tab.p,
                     op.elmat.p, op.elvec.p);
  hipLaunchKernelGGL(k_gather_q2, dim3((unsigned)((op.n + kQ2Rows - 1) / kQ2Rows)), dim3(64 * kQ2Rows), 0,
                     op.ctx->stream, t, op.n, s.d_rep.p, s.d_dofs.p, op.rowptr.p, op.col.p, op.elmat.p, op.elvec.p,
                     d_fixed, op.val.p, op.rhs.p);
}

template <int DIM, int ND, int NQ1, int MODE>
void launch_owner(femhip_operator &op, const PhysDev &ph, const double *d_fixed) {
  femhip_space &s = *op.space;
  if (op.n <= 0) return;
  // a row couples to the DOFs of the (2 p + 1)^d nodes around its own, per component
  int row_max = s.fe.ncomp;
  for (int d = 0; d < DIM; ++d) row_max *= 2 * s.fe.order + 1;
  FH_CHECK(row_max <= kOwnMaxRow, FEMHIP_ERR_INVALID, "row longer than the owner-computes kernel holds");
  if (MODE == 0 && op.trig.n == 0) build_trig(op);
  const TrigTab tt{reinterpret_cast<const double4 *>(op.trig.p), {op.trig_off[0], op.trig_off[1], op.trig_off[2]}};
  const double *tab = MODE == 2 ? op.elmat.p : op.tab.p;
  hipLaunchKernelGGL((k_assemble_owner<DIM, ND, NQ1, MODE>), dim3((unsigned)op.n), dim3(64), 0, op.ctx->stream, ph,
                     tt, face_tab(s), op.n, s.d_rep.p, s.d_dofs.p, op.rowptr.p, op.col.p, tab, d_fixed, op.val.p,
                     op.rhs.p, op.elvec.p);
}

// Cells with their own map, or coefficients at the quadrature points: every
// cell's element matrix from its own Jacobians (mapped.hip), then the rows'
// owners gather them in ascending cell order.
void launch_general(femhip_operator &op, const PhysDev &ph, const double *d_fixed) {
  femhip_space &s = *op.space;
  const RefFE &fe = s.fe;
  element_matrices_mapped(op);
  if (op.n <= 0) return;
  if (fe.dim == 3 && fe.nd == 27) {
    hipLaunchKernelGGL(k_gather_q2, dim3((unsigned)((op.n + kQ2Rows - 1) / kQ2Rows)), dim3(64 * kQ2Rows), 0,
                       op.ctx->stream, face_tab(s), op.n, s.d_rep.p, s.d_dofs.p, op.rowptr.p, op.col.p, op.elmat.p,
                       op.elvec.p, d_fixed, op.val.p, op.rhs.p);
  } else if (fe.dim == 2 && fe.nd == 4) {
    launch_owner<2, 4, 1, 2>(op, ph, d_fixed);
  } else if (fe.dim == 2 && fe.nd == 9) {
    launch_owner<2, 9, 1, 2>(op, ph, d_fixed);
  } else if (fe.dim == 3 && fe.nd == 8) {
    launch_owner<3, 8, 1, 2>(op, ph, d_fixed);
  } else if (fe.dim == 3 && fe.nd == 24) {
    launch_owner<3, 24, 1, 2>(op, ph, d_fixed);
  } else {
    throw Error{FEMHIP_ERR_INVALID, "unsupported element for per-cell integration"};
  }
}

VtxGrid vtx_grid(const femhip_mesh_desc_t &d) {
  VtxGrid g{};
  g.dim = d.dim;
  for (int k = 0; k < 3; ++k) {
    g.n[k] = k < d.dim ? d.n[k] : 1;
    g.nv[k] = k < d.dim ? d.n[k] + 1 : 1;
  }
  return g;
}

// Trig factors at the 1-D quadrature points of every cell layer: the d-th
// coordinate of a quadrature point is x_i (1 - xi) / 2 + x_{i+1} (1 + xi) / 2
// (the geometric Q1 map of sbm_cell_map.i90:673-687 along one direction).
// Layout: per direction, [layer][1-D point] of (sin pi x, cos pi x, sin 2 pi x,
// cos 2 pi x).
void build_trig(femhip_operator &op) {
  femhip_space &s = *op.space;
  const RefFE &fe = s.fe;
  std::vector<double> h;
  int stride = 1;   // index step of the quadrature points along d (first index fastest)
  for (int d = 0; d < fe.dim; ++d) {
    op.trig_off[d] = (int)(h.size() / 4);
    const int nd = s.mesh.n[d];
    const double lo = s.mesh.limits[2 * d], hi = s.mesh.limits[2 * d + 1];
    for (int i = 0; i < nd; ++i)
      for (int a = 0; a < fe.nq1; ++a) {
        const double xi = fe.qp[d][a * stride];   // quadrature point with 1-D index a along d, 0 elsewhere
        const double x0 = (hi - lo) * ((double)i / (double)nd) + lo;
        const double x1 = (hi - lo) * ((double)(i + 1) / (double)nd) + lo;
        const double x = x0 * (0.5 * (1.0 - xi)) + x1 * (0.5 * (1.0 + xi));
        h.push_back(std::sin(kPi * x));
        h.push_back(std::cos(kPi * x));
        h.push_back(std::sin(2 * kPi * x));
        h.push_back(std::cos(2 * kPi * x));
      }
    stride *= fe.nq1;
  }
  op.trig.alloc(h.size());
  FH_HIP(hipMemcpyAsync(op.trig.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, op.ctx->stream));
  FH_HIP(hipStreamSynchronize(op.ctx->stream));
}

// One-time set-up of the owner-computes path (scalar Q1 Poisson only); the
// maps are validated against the CSR pattern, and any mismatch keeps the
// operator on the coloured scatter.
void setup_gather(femhip_operator &op) {
  femhip_space &s = *op.space;
  const RefFE &fe = s.fe;
  const int dim = fe.dim;
  const bool off = getenv("FEMHIP_NO_GATHER") != nullptr;   // read per operator: tests route Q1 to k_assemble_owner
  if (off || op.tab_kind != 0 || fe.order != 1 || fe.ncomp != 1 || fe.nd != (1 << dim) || op.n <= 0) {
    op.gather = -1;
    return;
  }
  hipStream_t st = op.ctx->stream;
  const VtxGrid g = vtx_grid(s.mesh);
  const int64_t nverts = (int64_t)g.nv[0] * g.nv[1] * g.nv[2];
  const int ks = dim == 2 ? 16 : 32;
  op.vtx_dof.alloc((size_t)nverts);
  op.row_vtx.alloc((size_t)op.n);
  op.slot.alloc((size_t)op.n * ks);
  DevBuf<int> bad;
  bad.alloc(1);
  FH_HIP(hipMemsetAsync(op.row_vtx.p, 0xff, sizeof(int32_t) * (size_t)op.n, st));   // -1: no vertex
  FH_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), st));
  const int64_t t = s.ncells * fe.nd;
  const unsigned gb = (unsigned)((t + 255) / 256), rb = (unsigned)((op.n + 255) / 256);
  if (dim == 2) {
    hipLaunchKernelGGL(k_gather_vertices<2>, dim3(gb), dim3(256), 0, st, g, s.ncells, s.d_dofs.p, op.vtx_dof.p,
                       op.row_vtx.p);
    hipLaunchKernelGGL(k_gather_slots<2>, dim3(rb), dim3(256), 0, st, g, op.n, op.vtx_dof.p, op.row_vtx.p,
                       op.rowptr.p, op.col.p, op.slot.p, bad.p);
  } else {
    hipLaunchKernelGGL(k_gather_vertices<3>, dim3(gb), dim3(256), 0, st, g, s.ncells, s.d_dofs.p, op.vtx_dof.p,
                       op.row_vtx.p);
    hipLaunchKernelGGL(k_gather_slots<3>, dim3(rb), dim3(256), 0, st, g, op.n, op.vtx_dof.p, op.row_vtx.p,
                       op.rowptr.p, op.col.p, op.slot.p, bad.p);
  }
  FH_HIP(hipGetLastError());
  int h_bad = 1;
  FH_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  if (h_bad) {
    op.vtx_dof.release();
    op.row_vtx.release();
    op.slot.release();
    op.gather = -1;
    return;
  }
  build_trig(op);
  op.gather = 1;
}

constexpr int kGatherRows = 128;

}  // namespace

// The owner-computes rows can fill the symmetric-diagonal solver copy too when
// the column offset of each neighbour code is the same on every row (the
// structured numberings for which that copy exists).  The offsets are read off
// one row that has all its neighbours, mapped onto the stored diagonals (each
// used by exactly one code) and checked on every row on the device.
void setup_sym_gather(femhip_operator &op) {
  if (op.sym_gather != 0) return;
  op.sym_gather = -1;
  if (op.gather != 1 || !op.sl.sym || op.n <= 0) return;
  const int dim = op.space->fe.dim;
  const int K = dim == 2 ? 9 : 27, KS = dim == 2 ? 16 : 32;
  int32_t off[27] = {};
  bool found = false;
  for (const int64_t r : {(int64_t)op.n / 2, (int64_t)op.n / 3, 2 * (int64_t)op.n / 3}) {
    uint8_t sl[32];
    FH_HIP(hipMemcpy(sl, op.slot.p + r * KS, (size_t)KS, hipMemcpyDeviceToHost));
    bool full = true;
    for (int k = 0; k < K; ++k) full = full && sl[k] != 255;
    if (!full) continue;
    int32_t b = 0, cols[32];
    FH_HIP(hipMemcpy(&b, op.rowptr.p + r, sizeof(int32_t), hipMemcpyDeviceToHost));
    FH_HIP(hipMemcpy(cols, op.col.p + b, sizeof(int32_t) * (size_t)K, hipMemcpyDeviceToHost));
    for (int k = 0; k < K; ++k) off[k] = cols[sl[k]] - (int32_t)r;
    found = true;
    break;
  }
  if (!found) return;
  const std::vector<int32_t> &so = op.sl.h_soffs;
  uint64_t seen = 0;
  for (int k = 0; k < 32; ++k) op.sym_kidx[k] = -1;
  for (int k = 0; k < K; ++k) {
    if (off[k] < 0) continue;
    const auto it = std::find(so.begin(), so.end(), off[k]);
    if (it == so.end()) return;
    const int d = (int)(it - so.begin());
    if (d >= 64 || ((seen >> d) & 1)) return;
    seen |= 1ull << d;
    op.sym_kidx[k] = (int8_t)d;
  }
  if (so.size() > 64 || seen != (so.size() == 64 ? ~0ull : (1ull << so.size()) - 1)) return;
  CodeOffsets co{};
  for (int k = 0; k < K; ++k) co.off[k] = off[k];
  hipStream_t st = op.ctx->stream;
  DevBuf<int> bad;
  bad.alloc(1);
  FH_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), st));
  const unsigned rb = (unsigned)((op.n + 255) / 256);
  if (dim == 2)
    hipLaunchKernelGGL(k_check_code_offsets<2>, dim3(rb), dim3(256), 0, st, op.n, op.slot.p, op.rowptr.p, op.col.p,
                       co, bad.p);
  else
    hipLaunchKernelGGL(k_check_code_offsets<3>, dim3(rb), dim3(256), 0, st, op.n, op.slot.p, op.rowptr.p, op.col.p,
                       co, bad.p);
  FH_HIP(hipGetLastError());
  int h_bad = 1;
  FH_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  if (!h_bad) op.sym_gather = 1;
}

// FP32 symmetric-diagonal copy of A (sc == nullptr) or of S A S integrated on
// the matrix cores (k_assemble_sym32_mfma); false where that kernel does not
// apply (the copy is then derived from the FP64 one by the solver).
bool assemble_sym32(femhip_operator &op, const double *sc, float *v32) {
  if (op.gather != 1 || op.sym_gather != 1 || !op.sl.sym || op.tab_kind != 0 || op.n <= 0 || op.sl.kplus > 16 ||
      general_integration(op))
    return false;
  femhip_space &s = *op.space;
  const int dim = s.fe.dim;
  if (dim != 2 && dim != 3) return false;
  Sym32Out so{};
  so.v32 = v32;
  so.stride = op.sl.stride;
  so.nd = op.sl.kplus;
  const int K = dim == 2 ? 9 : 27;
  for (int d = 0; d < 16; ++d) {
    so.code[d] = -1;
    so.off[d] = d < op.sl.kplus ? op.sl.h_soffs[(size_t)d] : 0;
  }
  for (int k = 0; k < K; ++k)
    if (op.sym_kidx[k] >= 0) so.code[op.sym_kidx[k]] = (int8_t)k;
  const VtxGrid g = vtx_grid(s.mesh);
  const TrigTab tt{reinterpret_cast<const double4 *>(op.trig.p), {op.trig_off[0], op.trig_off[1], op.trig_off[2]}};
  const PhysDev ph{op.phys.kind, op.phys.k_kind, op.phys.f_kind, op.phys.k0, op.phys.f0, op.k_cell.p};
  const int64_t rows = op.sl.stride;   // padding rows [n, stride) are cleared
  const dim3 G((unsigned)((rows + kMfmaRows - 1) / kMfmaRows)), B(kMfmaRows);
  hipStream_t st = op.ctx->stream;
  if (dim == 2)
    hipLaunchKernelGGL(k_assemble_sym32_mfma<2>, G, B, 0, st, g, ph, tt, op.n, rows, op.row_vtx.p, op.slot.p,
                       op.tab.p, sc, so);
  else
    hipLaunchKernelGGL(k_assemble_sym32_mfma<3>, G, B, 0, st, g, ph, tt, op.n, rows, op.row_vtx.p, op.slot.p,
                       op.tab.p, sc, so);
  FH_HIP(hipGetLastError());
  return true;
}

namespace {

void launch_gather(femhip_operator &op, const PhysDev &ph, const double *d_fixed) {
  femhip_space &s = *op.space;
  hipStream_t st = op.ctx->stream;
  const VtxGrid g = vtx_grid(s.mesh);
  const TrigTab tt{reinterpret_cast<const double4 *>(op.trig.p), {op.trig_off[0], op.trig_off[1], op.trig_off[2]}};
  const unsigned rb = (unsigned)((op.n + kGatherRows - 1) / kGatherRows);
  // the solver's FP64 copy is written too (its padding rows stay as the first
  // fill left them); the FP32 copy is derived from it by the solver
  SymOut so{};
  op.sym_written = 0;
  if (op.sym_gather == 1 && op.sl.v64.p) {
    so.v64 = op.sl.v64.p;
    so.stride = op.sl.stride;
    so.kp = op.sl.kplus;
    for (int k = 0; k < 32; ++k) so.kidx[k] = op.sym_kidx[k];
    op.sym_written = 1;
  }
  // FEMHIP_ASSEMBLY_MFMA=1: 3-D rows on the matrix cores.  Not the default: on
  // gfx950 FP64 MFMA and FP64 vector FMA have the same peak, W is 8/27 dense and
  // the source chain uses 1 of 16 columns, so the MFMA kernel issues about 5x the
  // useful flops and took 0.91 ms against 0.44 ms at 128^3.  Read per call.
  const char *mf = getenv("FEMHIP_ASSEMBLY_MFMA");
  const bool mfma = mf && atoi(mf) == 1;
  if (s.fe.dim == 2)
    hipLaunchKernelGGL((k_assemble_gather<2, kGatherRows>), dim3(rb), dim3(kGatherRows), 0, st, g, ph, tt, op.n,
                       op.row_vtx.p, op.vtx_dof.p, op.slot.p, op.rowptr.p, op.tab.p, d_fixed, op.val.p, op.rhs.p, so);
  else if (mfma)
    hipLaunchKernelGGL(k_assemble_gather_mfma, dim3((unsigned)((op.n + kMfmaRows - 1) / kMfmaRows)), dim3(kMfmaRows),
                       0, st, g, ph, tt, op.n, op.row_vtx.p, op.vtx_dof.p, op.slot.p, op.rowptr.p, op.tab.p, d_fixed,
                       op.val.p, op.rhs.p, so);
  else
    hipLaunchKernelGGL((k_assemble_gather<3, kGatherRows>), dim3(rb), dim3(kGatherRows), 0, st, g, ph, tt, op.n,
                       op.row_vtx.p, op.vtx_dof.p, op.slot.p, op.rowptr.p, op.tab.p, d_fixed, op.val.p, op.rhs.p, so);
}

}  // namespace

// Element matrix (row-major nd x nd) and body-force vector of the irreducible
// elasticity plugin for a vector FE (sbm_irreducible_discrete_integration.i90:64-130):
// elmat(i,j) += dV grad(phi_i) : sigma(phi_j), sigma = lambda tr(eps) I + 2 mu eps,
// lambda, mu from E, nu as par_test_elasticity_constitutive_models.f90:34-37.
void elasticity_element(const RefFE &fe, const femhip_physics_t &ph, double *Ke, double *Fe) {
  const int nd = fe.nd, nq = fe.nq, nn = fe.nnode;
  const double E = ph.E, nu = ph.nu;
  const double lambda = (nu * E) / ((1 + nu) * (1 - 2 * nu));
  const double mu = E / (2 * (1 + nu));
  for (int i = 0; i < nd * nd; ++i) Ke[i] = 0.0;
  for (int i = 0; i < nd; ++i) Fe[i] = 0.0;
  for (int q = 0; q < nq; ++q) {
    double dV = fe.detj[q] * fe.qw[q];
    for (int i = 0; i < nd; ++i) {
      int ci = i / nn, ai = i % nn;
      for (int j = 0; j < nd; ++j) {
        int cj = j / nn, aj = j % nn;
        double gi[3], gj[3];
        for (int k = 0; k < 3; ++k) {
          gi[k] = fe.grad[ai][k][q];
          gj[k] = fe.grad[aj][k][q];
        }
        double divdiv = gi[ci] * gj[cj];
        double epseps = 0.5 * ((ci == cj ? gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2] : 0.0) + gi[cj] * gj[ci]);
        Ke[i * nd + j] += dV * (lambda * divdiv + 2 * mu * epseps);
      }
      Fe[i] += dV * ph.body[ci] * fe.shp[ai][q];
    }
  }
}

// Element tables for the operator's physics (host, once per set_physics).
void upload_physics(femhip_operator &op) {
  const RefFE &fe = op.space->fe;
  const femhip_physics_t &ph = op.phys;
  const int nd = fe.nd, nq = fe.nq, dim = fe.dim;
  std::vector<double> t;
  if (ph.kind == FEMHIP_PHYS_POISSON) {
    FH_CHECK(fe.ncomp == 1, FEMHIP_ERR_INVALID, "Poisson plugin needs a scalar FE space");
    t.assign((size_t)(nq * nd * nd + nq * nd), 0.0);
    for (int q = 0; q < nq; ++q) {
      double factor = fe.detj[q] * fe.qw[q];
      for (int i = 0; i < nd; ++i)
        for (int j = 0; j < nd; ++j) {
          double dot = 0.0;
          for (int k = 0; k < 3; ++k) dot += fe.grad[j][k][q] * fe.grad[i][k][q];
          t[(size_t)((q * nd + i) * nd + j)] = factor * dot;
        }
      for (int i = 0; i < nd; ++i) t[(size_t)(nq * nd * nd + q * nd + i)] = factor * fe.shp[i][q];
    }
    op.tab_kind = 0;
    upload_qpoint_values(op);
    if (ph.k_kind == FEMHIP_K_CELL) {
      FH_CHECK(ph.k_cell != nullptr, FEMHIP_ERR_INVALID, "k_cell is NULL for FEMHIP_K_CELL");
      op.k_cell.alloc((size_t)op.space->ncells);
      FH_HIP(hipMemcpy(op.k_cell.p, ph.k_cell, sizeof(double) * (size_t)op.space->ncells, hipMemcpyHostToDevice));
    }
  } else if (ph.kind == FEMHIP_PHYS_ELASTICITY) {
    FH_CHECK(fe.ncomp == dim && dim == 3, FEMHIP_ERR_INVALID, "elasticity plugin needs a 3-D vector FE space");
    t.assign((size_t)(nd * nd + nd), 0.0);
    elasticity_element(fe, ph, t.data(), t.data() + nd * nd);
    op.tab_kind = 1;
  } else {
    throw Error{FEMHIP_ERR_INVALID, "unknown physics kind"};
  }
  op.tab.alloc(t.size());
  FH_HIP(hipMemcpy(op.tab.p, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
}

void assemble(femhip_operator &op, const double *fixed_host, float *ms) {
  femhip_space &s = *op.space;
  hipStream_t st = op.ctx->stream;
  const double *d_fixed = nullptr;
  if (fixed_host && s.n_fixed > 0) {
    op.fixed.alloc((size_t)s.n_fixed);
    FH_HIP(hipMemcpyAsync(op.fixed.p, fixed_host, sizeof(double) * (size_t)s.n_fixed, hipMemcpyHostToDevice, st));
    d_fixed = op.fixed.p;
  }
  const RefFE &fe = s.fe;
  // the gather maps depend on the space only; whether they apply depends on the physics
  const bool general = general_integration(op);
  const bool poisson_q1 = !general && op.tab_kind == 0 && fe.order == 1 && fe.ncomp == 1 && fe.nd == (1 << fe.dim);
  if (poisson_q1 && op.gather == 0) setup_gather(op);
  const bool gather = poisson_q1 && op.gather == 1;
  op.sym_written = 0;
  hipEvent_t e0, e1;
  FH_HIP(hipEventCreate(&e0));
  FH_HIP(hipEventCreate(&e1));
  FH_HIP(hipEventRecord(e0, st));
  PhysDev ph{op.phys.kind, op.phys.k_kind, op.phys.f_kind, op.phys.k0, op.phys.f0, op.k_cell.p};
  // every value and RHS entry is written once by either path: no clearing pass
  if (general) {
    launch_general(op, ph, d_fixed);
  } else if (gather) {
    launch_gather(op, ph, d_fixed);
  } else if (op.tab_kind == 0) {
    if (fe.dim == 2 && fe.nd == 4) launch_owner<2, 4, 2, 0>(op, ph, d_fixed);
    else if (fe.dim == 2 && fe.nd == 9) launch_owner<2, 9, 3, 0>(op, ph, d_fixed);
    else if (fe.dim == 3 && fe.nd == 8) launch_owner<3, 8, 2, 0>(op, ph, d_fixed);
    else if (fe.dim == 3 && fe.nd == 27) launch_q2(op, ph, d_fixed);
    else throw Error{FEMHIP_ERR_INVALID, "unsupported Poisson element"};
  } else {
    if (fe.dim == 3 && fe.nd == 24) launch_owner<3, 24, 1, 1>(op, ph, d_fixed);
    else throw Error{FEMHIP_ERR_INVALID, "unsupported elasticity element"};
  }
  FH_HIP(hipGetLastError());
  FH_HIP(hipEventRecord(e1, st));
  FH_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  FH_HIP(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

void spmv_full(femhip_operator &op, const double *d_x, double *d_y) {
  if (op.n == 0) return;
  hipLaunchKernelGGL(k_spmv_csr, dim3((op.n + 255) / 256), dim3(256), 0, op.ctx->stream, op.n, op.rowptr.p,
                     op.col.p, op.val.p, d_x, d_y);
  FH_HIP(hipGetLastError());
}

}  // namespace femhip
