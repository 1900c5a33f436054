// multifield.hip -- several scalar fields on one structured mesh with a block
// layout, assembled into a block sparse matrix and solved by a block CG, on
// the device.
//
// FEMPAR's multi-field serial_fe_space_t numbers each field's DOFs by first
// touch with its own counter, started at the number of free DOFs of the
// earlier fields of the same block (sbm_serial_fe_space.i90:2676-2700), and
// its strong Dirichlet sweep runs field by field with one counter
// (:584-640).  Each field here is a femhip space numbered on the device as if
// alone; the block layout adds the offsets.  block_layout_t
// (Integration/FEOperator/block_layout.f90:36-264) maps fields to blocks and
// says which field pairs couple; the block_sparse_assembler_t
// (Integration/FEOperator/block_sparse_assembler.f90:47-224) inserts each
// coupled (row field, column field) part of the element matrix into block
// (block(f), block(g)) of a block_sparse_matrix_t
// (LinearAlgebra/SparseMatrix/block_sparse_matrix.f90), and the element
// vector into the block array.
//
// Device layout: per coupled block (ib, jb) a full-storage CSR (rowptr, col,
// val; a diagonal block's symmetric-storage view is taken row by row when the
// host asks), rows of a block ordered field by field; one vector of all
// blocks' rows concatenated in block order for the RHS and the solver.
// Kernels (one wavefront per block row): the pattern (the columns of the
// coupled fields in the cells around the row's vef, sorted in LDS), the
// assembly (the row's entries summed over those cells in ascending order from
// element tables, no atomics), and a CSR SpMV that adds the blocks of a block
// row in block order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "femhip_internal.h"

using namespace femhip;

namespace {

constexpr int kMaxFields = FEMHIP_MAX_FIELDS;
constexpr int kCand = 1024;   // 2^3 cells x 4 fields x 27 DOFs = 864 candidate columns at most

// The column side of one (row field, column block) launch: the coupled
// column fields, their DOF lists, local DOF counts, offsets in the block and
// the offset of the (row field, column field) element table.
struct ColSet {
  int32_t ng;
  const int32_t *dofs[kMaxFields];
  int32_t nd[kMaxFields];
  int32_t off[kMaxFields];
  int32_t tab[kMaxFields];
};

__device__ __forceinline__ int col_candidate(const ColSet &cs, int tot, int i, int &b, int &g, int &j) {
  b = i / tot;
  int rem = i % tot;
  g = 0;
  while (g < cs.ng - 1 && rem >= cs.nd[g]) {
    rem -= cs.nd[g];
    ++g;
  }
  j = rem;
  return 0;
}

// The cells around the vef of a row's representative (cell, local DOF):
// along each axis the vef is not tangent to, the two cells on either side
// (clipped to the mesh).  Bit e of b picks the upper cell along the e-th such
// axis, so b ascending visits the cells in ascending index.
struct Around {
  int ijk[3];
  int ax[3];
  int m;
  int f;     // the vef's n-face in the representative's cell
  int pos;   // the DOF's position in that n-face's own list
};

__device__ __forceinline__ Around around_of(const FaceTab &t, int64_t rep) {
  Around a;
  const int64_t cell = rep / t.nd;
  const int l = (int)(rep % t.nd);
  a.f = t.face_of[l];
  a.pos = t.pos_of[l];
  cell_ijk(t, cell, a.ijk);
  a.m = 0;
  for (int k = 0; k < t.dim; ++k)
    if (!((t.type[a.f] >> k) & 1)) a.ax[a.m++] = k;
  return a;
}

// cell b around the vef (false outside the mesh) and the row DOF's local
// index in it: the same vef is the n-face of the same type whose anchor is
// the vef's position relative to that cell
__device__ __forceinline__ bool around_cell(const FaceTab &t, const Around &a, int b, int64_t &cell, int &lc) {
  int cj[3] = {a.ijk[0], a.ijk[1], a.ijk[2]}, anc = 0;
  for (int e = 0; e < a.m; ++e) {
    const int k = a.ax[e], w = a.ijk[k] + ((t.anchor[a.f] >> k) & 1);
    cj[k] = w - 1 + ((b >> e) & 1);
    if (cj[k] < 0 || cj[k] >= t.n[k]) return false;
    anc |= (w - cj[k]) << k;
  }
  cell = cell_index(t, cj);
  lc = t.lst[t.ptr[t.face_by[t.type[a.f]][anc]] + a.pos];
  return true;
}

// Block pattern, one wavefront per row of the row field (block row = off +
// field row): the distinct free columns of the coupled column fields in the
// cells around the row's vef, sorted (bitonic, in LDS).  Pass 0 counts them
// (and those on or above the diagonal, for a diagonal block's symmetric
// storage), pass 1 writes them.
constexpr int kPatWaves = 2;
template <int PASS>
__global__ __launch_bounds__(kPatWaves * 64) void k_block_pattern(FaceTab t, int32_t nrows, int32_t row_off,
                                                                  const int64_t *__restrict__ rep, ColSet cs,
                                                                  int32_t *__restrict__ len,
                                                                  int32_t *__restrict__ ulen,
                                                                  const int32_t *__restrict__ rowptr,
                                                                  int32_t *__restrict__ col) {
  __shared__ int32_t buf[kPatWaves][kCand];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int32_t fr = blockIdx.x * kPatWaves + w;
  int32_t *s = buf[w];
  int tot = 0;
  for (int g = 0; g < cs.ng; ++g) tot += cs.nd[g];
  if (fr < nrows) {
    const Around a = around_of(t, rep[fr]);
    const int ncand = (1 << a.m) * tot;
    for (int q = 0; q < kCand / 64; ++q) {
      const int i = lane + 64 * q;
      int32_t v = INT_MAX;
      if (i < ncand) {
        int b, g, j;
        col_candidate(cs, tot, i, b, g, j);
        int64_t cell;
        int lc;
        if (around_cell(t, a, b, cell, lc)) {
          const int32_t d = cs.dofs[g][cell * cs.nd[g] + j];
          if (d >= 0) v = d + cs.off[g];
        }
      }
      s[i] = v;
    }
  } else {
    for (int q = 0; q < kCand / 64; ++q) s[lane + 64 * q] = INT_MAX;
  }
  __syncthreads();
  for (int size = 2; size <= kCand; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int q = 0; q < kCand / 128; ++q) {
        const int p = lane + 64 * q;
        const int i = 2 * p - (p & (stride - 1));
        const int j = i + stride;
        const bool up = (i & size) == 0;
        const int32_t x = s[i], y = s[j];
        if ((x > y) == up) {
          s[i] = y;
          s[j] = x;
        }
      }
      __syncthreads();
    }
  if (fr >= nrows) return;
  const int32_t r = row_off + fr;
  int32_t cnt = 0, ucnt = 0;
  const int32_t base = PASS == 1 ? rowptr[r] : 0;
  for (int q = 0; q < kCand / 64; ++q) {
    const int i = lane + 64 * q;
    const int32_t v = s[i];
    const bool first = v != INT_MAX && (i == 0 || s[i - 1] != v);
    const uint64_t mask = __ballot(first);
    if (PASS == 1 && first) col[base + cnt + __popcll(mask & ((1ull << lane) - 1))] = v;
    cnt += __popcll(mask);
    ucnt += __popcll(__ballot(first && v >= r));
  }
  if (PASS == 0 && lane == 0) {
    len[r] = cnt;
    ulen[r] = ucnt;
  }
}

// Block assembly, one wavefront per row of the row field: the row's columns
// and a zeroed value per column in LDS; then, cell by cell in ascending index
// and column field by column field, lane j adds the element entry (the row's
// local DOF in that cell, column DOF j) at the slot of its column (binary
// search); within one
// step the lanes' columns are distinct, so the sum of every entry runs over
// the cells in ascending order, the order of the reference's COO append +
// stable sort.  The row is written once.
__global__ __launch_bounds__(64) void k_block_assemble(FaceTab t, int32_t nrows, int32_t row_off,
                                                       const int64_t *__restrict__ rep, ColSet cs,
                                                       const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ col,
                                                       const double *__restrict__ tab, double *__restrict__ val) {
  __shared__ int32_t sc[kCand];
  __shared__ double sv[kCand];
  const int lane = threadIdx.x;
  const int32_t fr = blockIdx.x;
  if (fr >= nrows) return;
  const int32_t r = row_off + fr;
  const int32_t b0 = rowptr[r], len = rowptr[r + 1] - b0;
  for (int i = lane; i < len; i += 64) {
    sc[i] = col[b0 + i];
    sv[i] = 0.0;
  }
  __syncthreads();
  const Around a = around_of(t, rep[fr]);
  for (int b = 0; b < (1 << a.m); ++b) {
    int64_t cell;
    int l;
    if (!around_cell(t, a, b, cell, l)) continue;   // the same for the whole block (one row)
    for (int g = 0; g < cs.ng; ++g) {
      if (lane < cs.nd[g]) {
        const int32_t d = cs.dofs[g][cell * cs.nd[g] + lane];
        if (d >= 0) {
          const int32_t c = d + cs.off[g];
          int lo = 0;   // = the number of the row's columns below c (fixed steps, len <= kCand)
#pragma unroll
          for (int h = kCand / 2; h > 0; h >>= 1)
            if (lo + h <= len && sc[lo + h - 1] < c) lo += h;
          sv[lo] += tab[cs.tab[g] + l * cs.nd[g] + lane];
        }
      }
      __syncthreads();
    }
  }
  for (int i = lane; i < len; i += 64) val[b0 + i] = sv[i];
}

// Block RHS, one thread per row of the row field: the element vector entry of
// the row's local DOF summed over the cells around its vef in ascending order
// (element_serial_block_array_assembly adds cell by cell).
__global__ void k_block_rhs(FaceTab t, int32_t nrows, int32_t row_off, const int64_t *__restrict__ rep,
                            const double *__restrict__ elvec, double *__restrict__ rhs) {
  const int32_t fr = blockIdx.x * blockDim.x + threadIdx.x;
  if (fr >= nrows) return;
  const Around a = around_of(t, rep[fr]);
  double acc = 0.0;
  for (int b = 0; b < (1 << a.m); ++b) {
    int64_t cell;
    int l;
    if (around_cell(t, a, b, cell, l)) acc += elvec[l];
  }
  rhs[row_off + fr] = acc;
}

// y = A x (add = 0) or y += A x, full-storage CSR, LANES lanes per row
// (rows of 27 to 125 entries): the lanes stride over the row's entries, so
// the value and column loads of a row are contiguous across the lanes; each
// lane issues the loads of 4 of its entries before it uses the first, so a
// long row is not a chain of dependent round trips; a fixed shuffle tree sums
// the lanes; the blocks of a block row are applied in block order (a
// block_sparse_matrix_t apply).
constexpr int kSpmvThreads = 256;
template <int LANES>
__global__ __launch_bounds__(kSpmvThreads) void k_block_spmv(int32_t n, const int32_t *__restrict__ rowptr,
                                                             const int32_t *__restrict__ col,
                                                             const double *__restrict__ val,
                                                             const double *__restrict__ x, double *__restrict__ y,
                                                             int add, const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;   // the solve has ended (see k_cg_test)
  const int sub = threadIdx.x % LANES;
  const int64_t r = ((int64_t)blockIdx.x * kSpmvThreads + threadIdx.x) / LANES;
  double acc = 0.0;
  if (r < n) {
    int32_t k = rowptr[r] + sub;
    const int32_t e = rowptr[r + 1];
    for (; k + 3 * LANES < e; k += 4 * LANES) {
      int32_t c[4];
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        c[u] = col[k + u * LANES];
        v[u] = val[k + u * LANES];
      }
      double g[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) g[u] = x[c[u]];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc += v[u] * g[u];
    }
    for (; k < e; k += LANES) acc += val[k] * x[col[k]];
  }
  for (int o = LANES / 2; o > 0; o >>= 1) acc += __shfl_down(acc, o, LANES);
  if (r < n && sub == 0) y[r] = add ? y[r] + acc : acc;
}

// 16 lanes a row: 8 take the same time, 32 are 20 % slower (48^3 Q2 + Q1 block CG)
constexpr int kSpmvLanes = 16;

// dot products in a fixed order: kDotBlocks partial sums (grid-strided rows,
// block tree), then one block sums them in order
constexpr int kDotBlocks = 256, kDotThreads = 256;
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[kDotThreads];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kDotThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}
__global__ __launch_bounds__(kDotThreads) void k_dot_part(int64_t n, const double *__restrict__ a,
                                                          const double *__restrict__ b, double *__restrict__ part,
                                                          const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kDotThreads + threadIdx.x; i < n; i += (int64_t)kDotBlocks * kDotThreads)
    acc += a[i] * b[i];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ __launch_bounds__(kDotThreads) void k_dot_final(const double *__restrict__ part, double *__restrict__ out,
                                                           const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;
  const double s = block_sum(part[threadIdx.x]);
  if (threadIdx.x == 0) *out = s;
}

// Device scalars of the block CG: kRz = <r,z> of the current direction,
// kQp = <q,p>, kRzNew = the next <r,z>, kRr = ||r||^2; the stopping test
// (k_cg_test) keeps the iteration count, the estimate ||r||, whether it
// converged and kDone, which every kernel of an iteration reads at entry, so
// iterations queued after the end return at once.
enum { kRz = 0, kQp = 1, kRzNew = 2, kRr = 3, kDot = 4, kIt = 5, kEst = 6, kConv = 7, kDone = 8, kScalars = 9 };

// it += 1 ; est = ||r|| ; converged = est <= rhs_test (cg.f90:323-407, RES_RES
// and RES_RHS) ; done = converged or it >= maxit
__global__ void k_cg_test(double *__restrict__ sc, double rhs_test, int maxit) {
  if (sc[kDone] != 0.0) return;
  const double it = sc[kIt] + 1.0;
  const double est = sqrt(sc[kRr]);
  const bool conv = est <= rhs_test;
  sc[kIt] = it;
  sc[kEst] = est;
  sc[kConv] = conv ? 1.0 : 0.0;
  sc[kDone] = (conv || it >= (double)maxit) ? 1.0 : 0.0;
}
// a new solve: no iteration made, not done
__global__ void k_cg_reset(double *__restrict__ sc) {
  sc[kIt] = 0.0;
  sc[kEst] = 0.0;
  sc[kConv] = 0.0;
  sc[kDone] = 0.0;
}

// alpha = <r,z> / <q,p> (0 when <q,p> = 0, cg.f90:195-199); x += alpha p, r -= alpha q
__global__ void k_cg_xr(int64_t n, const double *__restrict__ sc, double *__restrict__ x,
                        const double *__restrict__ p, double *__restrict__ r, const double *__restrict__ q) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || sc[kDone] != 0.0) return;
  const double qp = sc[kQp];
  const double alpha = qp != 0.0 ? sc[kRz] / qp : 0.0;
  x[i] = x[i] + alpha * p[i];
  r[i] = r[i] + (-alpha) * q[i];
}
// z = invd o r (identity: z = r); p = beta p + z as scal then daxpy
__global__ void k_cg_z(int64_t n, const double *__restrict__ invd, const double *__restrict__ r,
                       double *__restrict__ z, const double *__restrict__ stop) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (stop && *stop != 0.0)) return;
  z[i] = invd ? invd[i] * r[i] : r[i];
}
// beta = (1 / <r,z>_old) <r,z>_new (cg.f90's order); p = beta p + z
__global__ void k_cg_p(int64_t n, const double *__restrict__ sc, const double *__restrict__ z,
                       double *__restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || sc[kDone] != 0.0) return;
  const double beta = (1.0 / sc[kRz]) * sc[kRzNew];
  double v = beta * p[i];
  p[i] = v + 1.0 * z[i];
}
// the new <r,z> becomes the current one (after every k_cg_p block read both)
__global__ void k_cg_shift(double *__restrict__ sc) {
  if (sc[kDone] == 0.0) sc[kRz] = sc[kRzNew];
}
// b - A x0 into r
__global__ void k_sub(int64_t n, const double *__restrict__ b, double *__restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - r[i];
}
// inverse of a diagonal block's diagonal (Jacobi on the block matrix); a
// zero pivot raises the flag
__global__ void k_block_invdiag(int32_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                const double *__restrict__ val, double *__restrict__ invd, int *__restrict__ bad) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double d = 0.0;
  for (int32_t k = rowptr[r]; k < rowptr[r + 1]; ++k)
    if (col[k] == r) d = val[k];
  if (d == 0.0) atomicOr(bad, 1);
  invd[r] = 1.0 / d;
}

unsigned nblocks(int64_t n, int t) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

}  // namespace

struct femhip_block_operator {
  femhip_ctx *ctx = nullptr;
  int nf = 0, nb = 0;
  femhip_space *field[kMaxFields] = {};
  int block[kMaxFields] = {};               // 0-based block of each field
  int coupling[kMaxFields][kMaxFields] = {};
  int sym_storage = 0;                      // diagonal blocks' symmetric storage (the host view)
  int32_t free_off[kMaxFields] = {};        // first row of each field in its block
  int32_t fixed_off[kMaxFields] = {};       // fixed ids of the earlier fields
  int32_t nblk[kMaxFields] = {};            // rows per block
  int64_t boff[kMaxFields + 1] = {};        // first row of each block in the block vectors
  struct Blk {
    bool on = false;
    DevBuf<int32_t> rowptr, col;
    DevBuf<double> val;
    int64_t nnz = 0, nnz_upper = 0;
  } blk[kMaxFields][kMaxFields];
  RefFE fe[kMaxFields];                      // each field on the common quadrature
  femhip_multifield_physics_t phys{};
  bool have_phys = false, computed = false;
  DevBuf<double> tab, elvec, rhs;
  int32_t tab_off[kMaxFields][kMaxFields] = {}, vec_off[kMaxFields] = {};
  DevBuf<double> x, r, z, p, q, invd, part, scal;
  int64_t ntot = 0;
  ~femhip_block_operator() {
    for (int f = 0; f < nf; ++f)
      if (field[f]) femhip_space_destroy(field[f]);
    if (ctx) ctx_release(ctx);
  }
};

namespace {

// the columns of block jb for rows of field f
ColSet col_set(const femhip_block_operator &B, int f, int jb) {
  ColSet cs{};
  for (int g = 0; g < B.nf; ++g) {
    if (B.block[g] != jb || !B.coupling[f][g]) continue;
    cs.dofs[cs.ng] = B.field[g]->d_dofs.p;
    cs.nd[cs.ng] = B.field[g]->fe.nd;
    cs.off[cs.ng] = B.free_off[g];
    cs.tab[cs.ng] = B.tab_off[f][g];
    ++cs.ng;
  }
  return cs;
}

void build_block_patterns(femhip_block_operator &B) {
  hipStream_t st = B.ctx->stream;
  for (int ib = 0; ib < B.nb; ++ib)
    for (int jb = 0; jb < B.nb; ++jb) {
      auto &K = B.blk[ib][jb];
      K.on = false;
      for (int f = 0; f < B.nf; ++f)
        for (int g = 0; g < B.nf; ++g) K.on = K.on || (B.block[f] == ib && B.block[g] == jb && B.coupling[f][g]);
      if (!K.on) continue;
      const int32_t n = B.nblk[ib];
      DevBuf<int32_t> len, ulen;
      len.alloc((size_t)n + 1);
      ulen.alloc((size_t)n + 1);
      FH_HIP(hipMemsetAsync(len.p, 0, sizeof(int32_t) * ((size_t)n + 1), st));
      FH_HIP(hipMemsetAsync(ulen.p, 0, sizeof(int32_t) * ((size_t)n + 1), st));
      for (int f = 0; f < B.nf; ++f) {
        if (B.block[f] != ib || B.field[f]->n_free == 0) continue;
        const ColSet cs = col_set(B, f, jb);
        if (cs.ng == 0) continue;
        hipLaunchKernelGGL(k_block_pattern<0>, dim3(nblocks(B.field[f]->n_free, kPatWaves)), dim3(kPatWaves * 64), 0,
                           st, face_tab(*B.field[f]), B.field[f]->n_free, B.free_off[f], B.field[f]->d_rep.p, cs,
                           len.p, ulen.p, nullptr, nullptr);
        FH_HIP(hipGetLastError());
      }
      K.rowptr.alloc((size_t)n + 1);
      K.nnz = exclusive_scan(st, len.p, K.rowptr.p, n, true);
      DevBuf<int32_t> utmp;
      utmp.alloc((size_t)n + 1);
      K.nnz_upper = exclusive_scan(st, ulen.p, utmp.p, n, false);
      K.col.alloc((size_t)std::max<int64_t>(1, K.nnz));
      K.val.alloc((size_t)std::max<int64_t>(1, K.nnz));
      for (int f = 0; f < B.nf; ++f) {
        if (B.block[f] != ib || B.field[f]->n_free == 0) continue;
        const ColSet cs = col_set(B, f, jb);
        if (cs.ng == 0) continue;
        hipLaunchKernelGGL(k_block_pattern<1>, dim3(nblocks(B.field[f]->n_free, kPatWaves)), dim3(kPatWaves * 64), 0,
                           st, face_tab(*B.field[f]), B.field[f]->n_free, B.free_off[f], B.field[f]->d_rep.p, cs,
                           nullptr, nullptr, K.rowptr.p, K.col.p);
        FH_HIP(hipGetLastError());
      }
      FH_HIP(hipStreamSynchronize(st));
    }
}

// Element tables of the coupled diffusion-reaction plugin for every field
// pair (row-major nd_f x nd_g) and element vectors, on the common quadrature,
// in the plugin's loop order (q outermost, then i, then j)
void upload_block_physics(femhip_block_operator &B) {
  const femhip_multifield_physics_t &ph = B.phys;
  std::vector<double> t, v;
  const RefFE &f0 = B.fe[0];
  for (int f = 0; f < B.nf; ++f)
    for (int g = 0; g < B.nf; ++g) {
      B.tab_off[f][g] = (int32_t)t.size();
      t.resize(t.size() + (size_t)B.fe[f].nd * B.fe[g].nd, 0.0);
    }
  for (int f = 0; f < B.nf; ++f) {
    B.vec_off[f] = (int32_t)v.size();
    v.resize(v.size() + (size_t)B.fe[f].nd, 0.0);
  }
  for (int q = 0; q < f0.nq; ++q) {
    const double dv = f0.detj[q] * f0.qw[q];
    for (int f = 0; f < B.nf; ++f) {
      const RefFE &A = B.fe[f];
      for (int i = 0; i < A.nd; ++i) {
        for (int g = 0; g < B.nf; ++g) {
          const RefFE &G = B.fe[g];
          for (int j = 0; j < G.nd; ++j) {
            double e;
            if (f == g) {
              double dot = 0.0;
              for (int k = 0; k < 3; ++k) dot = dot + A.grad[j][k][q] * A.grad[i][k][q];
              e = ph.kappa[f] * dot + (B.nf - 1) * ph.sigma * (A.shp[i][q] * A.shp[j][q]);
            } else {
              e = -ph.sigma * (A.shp[i][q] * G.shp[j][q]);
            }
            t[(size_t)B.tab_off[f][g] + (size_t)i * G.nd + j] += dv * e;
          }
        }
        v[(size_t)B.vec_off[f] + i] += dv * ph.f[f] * A.shp[i][q];
      }
    }
  }
  B.tab.alloc(t.size());
  FH_HIP(hipMemcpy(B.tab.p, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
  B.elvec.alloc(v.size());
  FH_HIP(hipMemcpy(B.elvec.p, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
}

void block_compute(femhip_block_operator &B, double *ms) {
  hipStream_t st = B.ctx->stream;
  hipEvent_t e0, e1;
  FH_HIP(hipEventCreate(&e0));
  FH_HIP(hipEventCreate(&e1));
  FH_HIP(hipEventRecord(e0, st));
  for (int ib = 0; ib < B.nb; ++ib)
    for (int jb = 0; jb < B.nb; ++jb) {
      auto &K = B.blk[ib][jb];
      if (!K.on) continue;
      for (int f = 0; f < B.nf; ++f) {
        if (B.block[f] != ib || B.field[f]->n_free == 0) continue;
        const ColSet cs = col_set(B, f, jb);
        if (cs.ng == 0) continue;
        hipLaunchKernelGGL(k_block_assemble, dim3((unsigned)B.field[f]->n_free), dim3(64), 0, st,
                           face_tab(*B.field[f]), B.field[f]->n_free, B.free_off[f], B.field[f]->d_rep.p, cs,
                           K.rowptr.p, K.col.p, B.tab.p, K.val.p);
        FH_HIP(hipGetLastError());
      }
    }
  for (int f = 0; f < B.nf; ++f) {
    if (B.field[f]->n_free == 0) continue;
    const int ib = B.block[f];
    hipLaunchKernelGGL(k_block_rhs, dim3(nblocks(B.field[f]->n_free, 256)), dim3(256), 0, st, face_tab(*B.field[f]),
                       B.field[f]->n_free, B.free_off[f], B.field[f]->d_rep.p, B.elvec.p + B.vec_off[f],
                       B.rhs.p + B.boff[ib]);
    FH_HIP(hipGetLastError());
  }
  FH_HIP(hipEventRecord(e1, st));
  FH_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  FH_HIP(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

// y = A x over the block vectors (device pointers)
void block_apply(femhip_block_operator &B, const double *x, double *y, const double *stop = nullptr) {
  hipStream_t st = B.ctx->stream;
  for (int ib = 0; ib < B.nb; ++ib) {
    int first = 1;
    const int32_t n = B.nblk[ib];
    if (n == 0) continue;
    for (int jb = 0; jb < B.nb; ++jb) {
      auto &K = B.blk[ib][jb];
      if (!K.on) continue;
      hipLaunchKernelGGL(k_block_spmv<kSpmvLanes>, dim3(nblocks((int64_t)n * kSpmvLanes, kSpmvThreads)),
                         dim3(kSpmvThreads), 0, st, n, K.rowptr.p, K.col.p, K.val.p, x + B.boff[jb], y + B.boff[ib],
                         first ? 0 : 1, stop);
      FH_HIP(hipGetLastError());
      first = 0;
    }
    if (first) FH_HIP(hipMemsetAsync(y + B.boff[ib], 0, sizeof(double) * (size_t)n, st));
  }
}

// <a, b> into the device scalar `slot` (no host wait)
void block_dot_dev(femhip_block_operator &B, const double *a, const double *b, int slot,
                   const double *stop = nullptr) {
  hipStream_t st = B.ctx->stream;
  hipLaunchKernelGGL(k_dot_part, dim3(kDotBlocks), dim3(kDotThreads), 0, st, B.ntot, a, b, B.part.p, stop);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kDotThreads), 0, st, B.part.p, B.scal.p + slot, stop);
  FH_HIP(hipGetLastError());
}

double read_scalar(femhip_block_operator &B, int slot) {
  double h = 0.0;
  FH_HIP(hipMemcpyAsync(&h, B.scal.p + slot, sizeof(double), hipMemcpyDeviceToHost, B.ctx->stream));
  FH_HIP(hipStreamSynchronize(B.ctx->stream));
  return h;
}

double block_dot(femhip_block_operator &B, const double *a, const double *b) {
  block_dot_dev(B, a, b, kDot);
  return read_scalar(B, kDot);
}

constexpr int kBlockChunk = 8;   // iterations queued between two reads of the device state

// cg_solve_body (cg.f90:101-238) on the block operator with the RES_RES or
// RES_RHS criterion (cg.f90:279-407), identity or Jacobi (the inverse
// diagonals of the diagonal blocks) preconditioner; x holds x0
void block_cg(femhip_block_operator &B, const femhip_cg_params_t &prm, const double *b, femhip_cg_result_t &res) {
  hipStream_t st = B.ctx->stream;
  const int64_t n = B.ntot;
  const dim3 G(nblocks(n, 256)), T(256);
  const bool jac = prm.preconditioner == FEMHIP_PREC_JACOBI;
  hipEvent_t e0, e1;
  FH_HIP(hipEventCreate(&e0));
  FH_HIP(hipEventCreate(&e1));
  FH_HIP(hipEventRecord(e0, st));
  if (jac) {
    DevBuf<int> bad;
    bad.alloc(1);
    FH_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), st));
    for (int ib = 0; ib < B.nb; ++ib) {
      const int32_t nr = B.nblk[ib];
      if (nr == 0) continue;
      FH_CHECK(B.blk[ib][ib].on, FEMHIP_ERR_ZERO_DIAGONAL, "diagonal block missing: Jacobi needs it");
      hipLaunchKernelGGL(k_block_invdiag, dim3(nblocks(nr, 256)), T, 0, st, nr, B.blk[ib][ib].rowptr.p,
                         B.blk[ib][ib].col.p, B.blk[ib][ib].val.p, B.invd.p + B.boff[ib], bad.p);
      FH_HIP(hipGetLastError());
    }
    int h_bad = 0;
    FH_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    FH_CHECK(!h_bad, FEMHIP_ERR_ZERO_DIAGONAL, "op1 cannot be entrywise inverted; there exists a null diagonal entry");
  }
  const double *invd = jac ? B.invd.p : nullptr;
  block_apply(B, B.x.p, B.r.p);
  hipLaunchKernelGGL(k_sub, G, T, 0, st, n, b, B.r.p);
  hipLaunchKernelGGL(k_cg_z, G, T, 0, st, n, invd, B.r.p, B.z.p, nullptr);
  FH_HIP(hipGetLastError());
  block_dot_dev(B, B.r.p, B.z.p, kRz);
  FH_HIP(hipMemcpyAsync(B.p.p, B.z.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, st));
  const double rhs_test =
      prm.stopping_criterion == FEMHIP_RES_RHS ? prm.rtol * std::sqrt(block_dot(B, b, b)) + prm.atol
                                               : prm.rtol * std::sqrt(block_dot(B, B.r.p, B.r.p)) + prm.atol;
  // alpha, beta and the stopping test stay on the device: the host queues
  // kBlockChunk iterations and reads the state once per chunk; the kernels of
  // iterations queued past the end return at entry (kDone)
  hipLaunchKernelGGL(k_cg_reset, dim3(1), dim3(1), 0, st, B.scal.p);
  const double *stop = B.scal.p + kDone;
  double h[4] = {0.0, 0.0, 0.0, 0.0};   // kIt, kEst, kConv, kDone
  static_assert(kEst == kIt + 1 && kConv == kIt + 2 && kDone == kIt + 3, "state slots are contiguous");
  for (;;) {
    for (int c = 0; c < kBlockChunk; ++c) {
      block_apply(B, B.p.p, B.q.p, stop);
      block_dot_dev(B, B.q.p, B.p.p, kQp, stop);
      hipLaunchKernelGGL(k_cg_xr, G, T, 0, st, n, B.scal.p, B.x.p, B.p.p, B.r.p, B.q.p);
      block_dot_dev(B, B.r.p, B.r.p, kRr, stop);
      hipLaunchKernelGGL(k_cg_test, dim3(1), dim3(1), 0, st, B.scal.p, rhs_test, prm.max_num_iterations);
      hipLaunchKernelGGL(k_cg_z, G, T, 0, st, n, invd, B.r.p, B.z.p, stop);
      block_dot_dev(B, B.r.p, B.z.p, kRzNew, stop);
      hipLaunchKernelGGL(k_cg_p, G, T, 0, st, n, B.scal.p, B.z.p, B.p.p);
      hipLaunchKernelGGL(k_cg_shift, dim3(1), dim3(1), 0, st, B.scal.p);
      FH_HIP(hipGetLastError());
    }
    FH_HIP(hipMemcpyAsync(h, B.scal.p + kIt, sizeof(h), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    if (h[3] != 0.0) break;
  }
  const int it = (int)h[0];
  const bool conv = h[2] != 0.0;
  const double est = h[1];
  FH_HIP(hipEventRecord(e1, st));
  FH_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  FH_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // true residual in FP64
  block_apply(B, B.x.p, B.q.p);
  hipLaunchKernelGGL(k_sub, G, T, 0, st, n, b, B.q.p);
  res.true_residual = std::sqrt(block_dot(B, B.q.p, B.q.p));
  res.rhs_norm = std::sqrt(block_dot(B, b, b));
  res.num_iterations = it;
  res.did_converge = conv ? 1 : 0;
  res.outer_iterations = 0;
  res.status = conv ? FEMHIP_OK : FEMHIP_ERR_NOT_CONVERGED;
  res.error_estimate = est;
  res.rhs_convergence_test = rhs_test;
  res.solve_ms = t;
}

void check_block(const femhip_block_operator *B, int32_t ib, int32_t jb) {
  FH_CHECK(B, FEMHIP_ERR_INVALID, "null block operator");
  FH_CHECK(ib >= 1 && ib <= B->nb && jb >= 1 && jb <= B->nb, FEMHIP_ERR_INVALID, "block index out of range");
}

}  // namespace

extern "C" {

int femhip_block_operator_create(femhip_ctx_t *ctx, const femhip_block_layout_t *layout,
                                 int32_t diagonal_blocks_symmetric_storage, femhip_block_operator_t **bop) {
  try {
    FH_CHECK(ctx && layout && bop, FEMHIP_ERR_INVALID, "null argument");
    const int nf = layout->num_fields;
    FH_CHECK(nf >= 1 && nf <= kMaxFields, FEMHIP_ERR_INVALID, "num_fields must be in [1, 4]");
    int maxo = 0;
    for (int f = 0; f < nf; ++f) {
      const femhip_space *s = layout->fields[f];
      FH_CHECK(s, FEMHIP_ERR_INVALID, "null field space");
      FH_CHECK(s->ctx == ctx, FEMHIP_ERR_INVALID, "field spaces must belong to the context");
      FH_CHECK(s->fe.ncomp == 1, FEMHIP_ERR_INVALID, "block operator fields must be scalar");
      const femhip_mesh_desc_t &a = s->mesh, &b = layout->fields[0]->mesh;
      bool same = a.dim == b.dim;
      for (int k = 0; k < 3; ++k) same = same && a.n[k] == b.n[k];
      for (int k = 0; k < 6; ++k) same = same && a.limits[k] == b.limits[k];
      FH_CHECK(same, FEMHIP_ERR_INVALID, "all fields must live on the same mesh");
      FH_CHECK(layout->field_blocks[f] >= 1 && layout->field_blocks[f] <= nf, FEMHIP_ERR_INVALID,
               "field_blocks entries must be in [1, num_fields]");
      FH_CHECK(layout->field_coupling[f * nf + f] != 0, FEMHIP_ERR_INVALID, "every field must couple to itself");
      maxo = std::max(maxo, s->fe.order);
    }
    FH_HIP(hipSetDevice(ctx->device));
    auto *B = new femhip_block_operator();
    try {
      B->ctx = ctx_retain(ctx);
      B->nf = nf;
      B->sym_storage = diagonal_blocks_symmetric_storage ? 1 : 0;
      for (int f = 0; f < nf; ++f) {
        B->field[f] = layout->fields[f];
        B->field[f]->refs.fetch_add(1);
        B->block[f] = layout->field_blocks[f] - 1;
        B->nb = std::max(B->nb, B->block[f] + 1);
        for (int g = 0; g < nf; ++g) B->coupling[f][g] = layout->field_coupling[f * nf + g] != 0;
      }
      // block_layout_t%create: block coupling of coupled fields; offsets in field order
      int32_t fixed = 0;
      for (int f = 0; f < nf; ++f) {
        B->free_off[f] = B->nblk[B->block[f]];
        int64_t nb = (int64_t)B->nblk[B->block[f]] + B->field[f]->n_free;
        FH_CHECK(nb < INT32_MAX, FEMHIP_ERR_INVALID, "block too large for 32-bit ids");
        B->nblk[B->block[f]] = (int32_t)nb;
        B->fixed_off[f] = fixed;
        fixed += B->field[f]->n_fixed;
      }
      for (int b = 0; b < B->nb; ++b) {
        FH_CHECK(B->nblk[b] >= 0, FEMHIP_ERR_INVALID, "bad block");
        B->boff[b + 1] = B->boff[b] + B->nblk[b];
      }
      B->ntot = B->boff[B->nb];
      // every field on the quadrature of the highest order (sbm_serial_fe_space.i90:1891-2006)
      const femhip_mesh_desc_t &m = B->field[0]->mesh;
      double h[3] = {1.0, 1.0, 1.0};
      for (int k = 0; k < m.dim; ++k) {
        const double a = m.limits[2 * k], b = m.limits[2 * k + 1];
        h[k] = ((b - a) * (1.0 / (double)m.n[k]) + a) - a;
      }
      for (int f = 0; f < nf; ++f) build_reffe(B->fe[f], m.dim, B->field[f]->fe.order, 1, h, maxo + 1);
      build_block_patterns(*B);
      const size_t nt = (size_t)std::max<int64_t>(1, B->ntot);
      B->rhs.alloc(nt);
      B->x.alloc(nt);
      B->r.alloc(nt);
      B->z.alloc(nt);
      B->p.alloc(nt);
      B->q.alloc(nt);
      B->invd.alloc(nt);
      B->part.alloc(kDotBlocks);
      B->scal.alloc(kScalars);
    } catch (...) {
      delete B;
      throw;
    }
    *bop = B;
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return FEMHIP_ERR_OOM;
  }
}

int femhip_block_operator_info(const femhip_block_operator_t *B, int32_t *num_blocks, int32_t *block_sizes,
                               int64_t *nnz) {
  try {
    FH_CHECK(B && num_blocks, FEMHIP_ERR_INVALID, "null argument");
    *num_blocks = B->nb;
    for (int ib = 0; ib < B->nb; ++ib) {
      if (block_sizes) block_sizes[ib] = B->nblk[ib];
      for (int jb = 0; jb < B->nb; ++jb) {
        const auto &K = B->blk[ib][jb];
        if (nnz) nnz[ib * B->nb + jb] = !K.on ? 0 : (B->sym_storage && ib == jb ? K.nnz_upper : K.nnz);
      }
    }
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_block_operator_get_dofs(const femhip_block_operator_t *B, int32_t field, int32_t *lst_dofs_gids) {
  try {
    FH_CHECK(B && lst_dofs_gids, FEMHIP_ERR_INVALID, "null argument");
    FH_CHECK(field >= 1 && field <= B->nf, FEMHIP_ERR_INVALID, "field index out of range");
    const femhip_space &s = *B->field[field - 1];
    ensure_host_dofs(s);
    const int32_t fo = B->free_off[field - 1], xo = B->fixed_off[field - 1];
    for (size_t i = 0; i < s.dofs.size(); ++i) {
      const int32_t g = s.dofs[i];
      lst_dofs_gids[i] = g > 0 ? g + fo : g - xo;
    }
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_block_operator_get_block_pattern(const femhip_block_operator_t *B, int32_t ib, int32_t jb, int32_t *irp,
                                            int32_t *ja) {
  try {
    check_block(B, ib, jb);
    FH_CHECK(irp && ja, FEMHIP_ERR_INVALID, "null argument");
    const auto &K = B->blk[ib - 1][jb - 1];
    const int32_t n = B->nblk[ib - 1];
    if (!K.on) {
      for (int32_t r = 0; r <= n; ++r) irp[r] = 1;
      return FEMHIP_OK;
    }
    FH_HIP(hipSetDevice(B->ctx->device));
    std::vector<int32_t> rp((size_t)n + 1), cl((size_t)std::max<int64_t>(1, K.nnz));
    FH_HIP(hipMemcpy(rp.data(), K.rowptr.p, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost));
    FH_HIP(hipMemcpy(cl.data(), K.col.p, sizeof(int32_t) * (size_t)K.nnz, hipMemcpyDeviceToHost));
    const bool up = B->sym_storage && ib == jb;
    int64_t k = 0;
    irp[0] = 1;
    for (int32_t r = 0; r < n; ++r) {
      for (int32_t e = rp[r]; e < rp[r + 1]; ++e)
        if (!up || cl[e] >= r) ja[k++] = cl[e] + 1;
      irp[r + 1] = (int32_t)(k + 1);
    }
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_block_operator_set_physics(femhip_block_operator_t *B, const femhip_multifield_physics_t *ph) {
  try {
    FH_CHECK(B && ph, FEMHIP_ERR_INVALID, "null argument");
    FH_CHECK(ph->sigma >= 0.0, FEMHIP_ERR_INVALID, "sigma must be >= 0");
    FH_HIP(hipSetDevice(B->ctx->device));
    B->phys = *ph;
    upload_block_physics(*B);
    B->have_phys = true;
    B->computed = false;
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_block_operator_compute(femhip_block_operator_t *B, double *ms) {
  try {
    FH_CHECK(B, FEMHIP_ERR_INVALID, "null block operator");
    FH_CHECK(B->have_phys, FEMHIP_ERR_STATE, "set the physics before compute");
    FH_HIP(hipSetDevice(B->ctx->device));
    block_compute(*B, ms);
    B->computed = true;
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_block_operator_get_block_values(const femhip_block_operator_t *B, int32_t ib, int32_t jb, double *val) {
  try {
    check_block(B, ib, jb);
    FH_CHECK(val, FEMHIP_ERR_INVALID, "null argument");
    FH_CHECK(B->computed, FEMHIP_ERR_STATE, "block operator not computed");
    const auto &K = B->blk[ib - 1][jb - 1];
    if (!K.on) return FEMHIP_OK;
    FH_HIP(hipSetDevice(B->ctx->device));
    const int32_t n = B->nblk[ib - 1];
    std::vector<int32_t> rp((size_t)n + 1), cl((size_t)std::max<int64_t>(1, K.nnz));
    std::vector<double> v((size_t)std::max<int64_t>(1, K.nnz));
    FH_HIP(hipMemcpy(rp.data(), K.rowptr.p, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost));
    FH_HIP(hipMemcpy(cl.data(), K.col.p, sizeof(int32_t) * (size_t)K.nnz, hipMemcpyDeviceToHost));
    FH_HIP(hipMemcpy(v.data(), K.val.p, sizeof(double) * (size_t)K.nnz, hipMemcpyDeviceToHost));
    const bool up = B->sym_storage && ib == jb;
    int64_t k = 0;
    for (int32_t r = 0; r < n; ++r)
      for (int32_t e = rp[r]; e < rp[r + 1]; ++e)
        if (!up || cl[e] >= r) val[k++] = v[e];
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_block_operator_get_rhs(const femhip_block_operator_t *B, int32_t ib, double *rhs) {
  try {
    check_block(B, ib, ib);
    FH_CHECK(rhs, FEMHIP_ERR_INVALID, "null argument");
    FH_CHECK(B->computed, FEMHIP_ERR_STATE, "block operator not computed");
    FH_HIP(hipSetDevice(B->ctx->device));
    if (B->nblk[ib - 1] > 0)
      FH_HIP(hipMemcpy(rhs, B->rhs.p + B->boff[ib - 1], sizeof(double) * (size_t)B->nblk[ib - 1],
                       hipMemcpyDeviceToHost));
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_block_operator_apply(femhip_block_operator_t *B, const double *x, double *y) {
  try {
    FH_CHECK(B && x && y, FEMHIP_ERR_INVALID, "null argument");
    FH_CHECK(B->computed, FEMHIP_ERR_STATE, "block operator not computed");
    FH_HIP(hipSetDevice(B->ctx->device));
    if (B->ntot == 0) return FEMHIP_OK;
    const size_t bytes = sizeof(double) * (size_t)B->ntot;
    hipStream_t st = B->ctx->stream;
    FH_HIP(hipMemcpyAsync(B->z.p, x, bytes, hipMemcpyHostToDevice, st));
    block_apply(*B, B->z.p, B->q.p);
    FH_HIP(hipMemcpyAsync(y, B->q.p, bytes, hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_block_operator_solve(femhip_block_operator_t *B, const femhip_cg_params_t *prm, const double *b,
                                double *x, femhip_cg_result_t *res) {
  try {
    FH_CHECK(B && prm && x && res, FEMHIP_ERR_INVALID, "null argument");
    FH_CHECK(B->computed, FEMHIP_ERR_STATE, "block operator not computed");
    FH_CHECK(prm->precision == FEMHIP_PRECISION_FP64, FEMHIP_ERR_INVALID, "the block solver runs in FP64");
    FH_CHECK(prm->stopping_criterion == FEMHIP_RES_RES || prm->stopping_criterion == FEMHIP_RES_RHS,
             FEMHIP_ERR_INVALID, "the block solver supports the RES_RES and RES_RHS criteria");
    FH_CHECK(prm->preconditioner == FEMHIP_PREC_IDENTITY || prm->preconditioner == FEMHIP_PREC_JACOBI,
             FEMHIP_ERR_INVALID, "unknown preconditioner");
    FH_CHECK(prm->max_num_iterations >= 1 && prm->rtol >= 0 && prm->atol >= 0, FEMHIP_ERR_INVALID,
             "bad solver parameters");
    FH_HIP(hipSetDevice(B->ctx->device));
    *res = femhip_cg_result_t{};
    if (B->ntot == 0) {
      res->did_converge = 1;
      return FEMHIP_OK;
    }
    const size_t bytes = sizeof(double) * (size_t)B->ntot;
    hipStream_t st = B->ctx->stream;
    const double *db = B->rhs.p;
    DevBuf<double> bb;
    if (b) {
      bb.alloc((size_t)B->ntot);
      FH_HIP(hipMemcpyAsync(bb.p, b, bytes, hipMemcpyHostToDevice, st));
      db = bb.p;
    }
    FH_HIP(hipMemcpyAsync(B->x.p, x, bytes, hipMemcpyHostToDevice, st));
    block_cg(*B, *prm, db, *res);
    FH_HIP(hipMemcpyAsync(x, B->x.p, bytes, hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}

int femhip_block_operator_destroy(femhip_block_operator_t *B) {
  if (!B) return FEMHIP_OK;
  (void)hipSetDevice(B->ctx->device);
  (void)hipStreamSynchronize(B->ctx->stream);
  delete B;
  return FEMHIP_OK;
}

}  // extern "C"
