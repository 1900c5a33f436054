// setup.hip -- device-side set-up of a structured FE space and of its
// operator's CSR pattern: first-touch DOF numbering, the symbolic pattern,
// the symmetric-diagonal layout test and the strong Dirichlet values.
//
// First touch on a structured mesh in closed form.  The reference numbers DOFs
// in two sweeps over the cells in index order (strong Dirichlet DOFs -1, -2, ...
// first, then free DOFs 1, 2, ...; within a cell the interior first and then the
// n-faces in polytope order), and a vef's DOFs are numbered by the first cell
// that touches it (Integration/FESpace/sbm_serial_fe_space.i90:525-715,
// 2623-2922).  Whether a vef is Dirichlet depends on its position only, and the
// cell index grows with each coordinate, so the first cell touching a vef is the
// one with the smallest coordinate along every axis the vef is not tangent to:
// cell o numbers the vefs of its faces whose anchor-0 axes have o_k = 0.  Each
// cell's count of new DOFs is then local, an exclusive scan over the cells gives
// every cell's first id, and a DOF's id is its owner's first id plus the DOFs
// the owner numbers before it.  Every step is a per-cell or per-row kernel.
#include <algorithm>
#include <climits>

#include "femhip_internal.h"

namespace femhip {

FaceTab face_tab(const femhip_space &s) {
  const RefFE &fe = s.fe;
  FaceTab t{};
  t.dim = fe.dim;
  t.nd = fe.nd;
  t.nnf = fe.nnf;
  t.fc = fe.nnf - 1;
  t.bc = s.mesh.dirichlet;
  t.np1 = fe.np1;
  t.nnode = fe.nnode;
  t.order = fe.order;
  for (int k = 0; k < 3; ++k) {
    t.n[k] = k < fe.dim ? s.mesh.n[k] : 1;
    t.lo[k] = s.mesh.limits[2 * k];
    t.hi[k] = s.mesh.limits[2 * k + 1];
  }
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b) t.face_by[a][b] = -1;
  for (int f = 0; f < fe.nnf; ++f) {
    t.type[f] = fe.nf_type[f];
    t.anchor[f] = fe.nf_anchor[f];
    t.ptr[f] = fe.own_ptr[f];
    t.cnt[f] = fe.own_ptr[f + 1] - fe.own_ptr[f];
    t.face_by[fe.nf_type[f]][fe.nf_anchor[f]] = (int8_t)f;
    for (int k = fe.own_ptr[f]; k < fe.own_ptr[f + 1]; ++k) {
      t.face_of[fe.own_lst[k]] = (int8_t)f;
      t.pos_of[fe.own_lst[k]] = (int8_t)(k - fe.own_ptr[f]);
    }
  }
  t.ptr[fe.nnf] = fe.own_ptr[fe.nnf];
  for (int k = 0; k < fe.own_ptr[fe.nnf]; ++k) t.lst[k] = (int8_t)fe.own_lst[k];
  return t;
}

namespace {

// the vef of n-face f of cell ijk carries the strong Dirichlet condition
// (test_poisson_driver.f90:193-204: set_id 1 on the boundary; or the x = x_min plane)
__device__ __forceinline__ bool face_dirichlet(const FaceTab &t, const int (&ijk)[3], int f) {
  if (t.bc == FEMHIP_BC_NONE) return false;
  for (int k = 0; k < t.dim; ++k) {
    if ((t.type[f] >> k) & 1) continue;
    const int v = ijk[k] + ((t.anchor[f] >> k) & 1);
    if (t.bc == FEMHIP_BC_ALL_BOUNDARY && (v == 0 || v == t.n[k])) return true;
    if (t.bc == FEMHIP_BC_X0_FACE && k == 0 && v == 0) return true;
  }
  return false;
}

// no cell before ijk touches the vef of its n-face f
__device__ __forceinline__ bool owns(const FaceTab &t, const int (&ijk)[3], int f) {
  for (int k = 0; k < t.dim; ++k)
    if (!((t.type[f] >> k) & 1) && !((t.anchor[f] >> k) & 1) && ijk[k] > 0) return false;
  return true;
}

// DOFs cell ijk numbers in a sweep before its n-face `before` (fixed sweep:
// its Dirichlet vefs; free sweep: its interior, then its free vefs)
__device__ __forceinline__ int numbered_before(const FaceTab &t, const int (&ijk)[3], int before, bool fixed) {
  int p = fixed ? 0 : t.cnt[t.fc];
  for (int f = 0; f < before; ++f)
    if (t.cnt[f] > 0 && owns(t, ijk, f) && face_dirichlet(t, ijk, f) == fixed) p += t.cnt[f];
  return p;
}

__global__ void k_count_new(FaceTab t, int64_t ncells, int32_t *__restrict__ nfree, int32_t *__restrict__ nfixed) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  int ijk[3];
  cell_ijk(t, c, ijk);
  nfree[c] = numbered_before(t, ijk, t.fc, false);
  nfixed[c] = numbered_before(t, ijk, t.fc, true);
}

// dofs[c nd + l]: free -> 0-based row, fixed -> -(1-based fixed id); rep[row]
// = c nd + l of the DOF's owner cell (the row's representative)
__global__ void k_number(FaceTab t, int64_t ncells, const int32_t *__restrict__ base_free,
                         const int32_t *__restrict__ base_fixed, int32_t *__restrict__ dofs,
                         int64_t *__restrict__ rep) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  int ijk[3];
  cell_ijk(t, c, ijk);
  for (int f = 0; f < t.nnf; ++f) {
    if (t.cnt[f] == 0) continue;
    bool fixed = false;
    int64_t o = c;
    int pre = 0;
    if (f != t.fc) {
      fixed = face_dirichlet(t, ijk, f);
      int oijk[3] = {ijk[0], ijk[1], ijk[2]}, a = 0;
      for (int k = 0; k < t.dim; ++k) {
        if ((t.type[f] >> k) & 1) continue;
        const int w = ijk[k] + ((t.anchor[f] >> k) & 1);
        oijk[k] = w > 0 ? w - 1 : 0;
        a |= (w - oijk[k]) << k;
      }
      const int fs = t.face_by[t.type[f]][a];
      o = cell_index(t, oijk);
      pre = numbered_before(t, oijk, fs, fixed);
    }
    const int32_t base = (fixed ? base_fixed[o] : base_free[o]) + pre;
    for (int k = 0; k < t.cnt[f]; ++k) {
      const int l = t.lst[t.ptr[f] + k];
      const int32_t id = base + k + 1;
      dofs[c * t.nd + l] = fixed ? -id : id - 1;
      if (!fixed && o == c) rep[id - 1] = c * t.nd + l;
    }
  }
}

// ------------------------------------------------------------------ scan ---
// Exclusive scan of int32 counts in three launches: block sums, one block
// scanning them (int64, so an overflowing total is seen), block offsets added.
constexpr int kScanThreads = 256, kScanItems = 4, kScanBlock = kScanThreads * kScanItems;

__device__ __forceinline__ int64_t block_exclusive(int64_t v, int64_t *total) {
  __shared__ int64_t ws[kScanThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t inc = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) ws[w] = inc;
  __syncthreads();
  int64_t before = 0, all = 0;
  for (int i = 0; i < kScanThreads / 64; ++i) {
    if (i < w) before += ws[i];
    all += ws[i];
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_local(const int32_t *__restrict__ in, int32_t *__restrict__ out,
                                                             int64_t *__restrict__ bsum, int64_t n) {
  const int64_t i0 = (int64_t)blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  int32_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = i0 + k < n ? in[i0 + k] : 0;
    s += v[k];
  }
  int64_t tot;
  int64_t run = block_exclusive(s, &tot);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (i0 + k < n) out[i0 + k] = (int32_t)run;   // block-local; k_scan_add adds the block offset
    run += v[k];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_sums(int64_t *__restrict__ bsum, int64_t nb) {
  int64_t carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += kScanThreads) {
    const int64_t i = b0 + threadIdx.x;
    const int64_t v = i < nb ? bsum[i] : 0;
    int64_t tot;
    const int64_t ex = block_exclusive(v, &tot);
    if (i < nb) bsum[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(int32_t *__restrict__ out, const int64_t *__restrict__ bsum,
                                                           int64_t n) {
  const int64_t i0 = (int64_t)blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
  const int64_t b = bsum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (i0 + k < n) out[i0 + k] = (int32_t)(out[i0 + k] + b);
}

}  // namespace

// out[i] = sum of in[0..i) for i <= n (out holds n + 1 entries when with_total);
// returns the total
int64_t exclusive_scan(hipStream_t st, const int32_t *in, int32_t *out, int64_t n, bool with_total) {
  const int64_t nb = std::max<int64_t>(1, (n + kScanBlock - 1) / kScanBlock);
  DevBuf<int64_t> bsum;
  bsum.alloc((size_t)nb + 1);
  if (n > 0) {
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, out, bsum.p, n);
    FH_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, st, bsum.p, n > 0 ? nb : 0);
  FH_HIP(hipGetLastError());
  if (n > 0) {
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kScanThreads), 0, st, out, bsum.p, n);
    FH_HIP(hipGetLastError());
  }
  int64_t total = 0;
  FH_HIP(hipMemcpyAsync(&total, bsum.p + (n > 0 ? nb : 0), sizeof(int64_t), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  FH_CHECK(total < (int64_t)INT32_MAX, FEMHIP_ERR_INVALID, "count exceeds 2^31-1 (32-bit ids and CSR offsets)");
  if (with_total) FH_HIP(hipMemcpy(out + n, &total, sizeof(int32_t), hipMemcpyHostToDevice));
  return total;
}

namespace {

// --------------------------------------------------------------- pattern ---
// One wavefront per row.  The row's DOF sits on the vef of its representative
// (cell, local DOF); the cells around that vef (up to 2^d) hold every DOF it
// couples to.  Their free DOFs (at most 2^d nd <= 256) are sorted in LDS
// (bitonic) and the distinct ones counted (pass 0: row length and entries on or
// above the diagonal) or written in order (pass 1).  The pattern is the sorted
// set of distinct (row, col) pairs of free DOFs sharing a cell, what the
// reference's COO append + sort_and_compress yields
// (base_sparse_matrix.f90:3499-3510, 4162-4479).
constexpr int kPatWaves = 4, kPatCand = 256;

template <int PASS>
__global__ __launch_bounds__(kPatWaves * 64) void k_pattern(FaceTab t, int32_t n, const int64_t *__restrict__ rep,
                                                            const int32_t *__restrict__ dofs,
                                                            int32_t *__restrict__ len, int32_t *__restrict__ ulen,
                                                            const int32_t *__restrict__ rowptr,
                                                            int32_t *__restrict__ col) {
  __shared__ int32_t a[kPatWaves][kPatCand];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int32_t r = blockIdx.x * kPatWaves + w;
  int32_t *s = a[w];
  if (r < n) {
    const int64_t c = rep[r];
    const int64_t cell = c / t.nd;
    const int l = (int)(c % t.nd), f = t.face_of[l];
    int ijk[3];
    cell_ijk(t, cell, ijk);
    int ax[3], m = 0;
    for (int k = 0; k < t.dim; ++k)
      if (!((t.type[f] >> k) & 1)) ax[m++] = k;
    const int ncand = (1 << m) * t.nd;
    for (int q = 0; q < kPatCand / 64; ++q) {
      const int i = lane + 64 * q;
      int32_t v = INT_MAX;
      if (i < ncand) {
        const int b = i / t.nd, li = i % t.nd;
        int cj[3] = {ijk[0], ijk[1], ijk[2]};
        bool ok = true;
        for (int e = 0; e < m; ++e) {
          const int k = ax[e];
          cj[k] = ijk[k] + ((t.anchor[f] >> k) & 1) - 1 + ((b >> e) & 1);
          ok = ok && cj[k] >= 0 && cj[k] < t.n[k];
        }
        if (ok) {
          const int32_t d = dofs[cell_index(t, cj) * t.nd + li];
          if (d >= 0) v = d;
        }
      }
      s[i] = v;
    }
  } else {
    for (int q = 0; q < kPatCand / 64; ++q) s[lane + 64 * q] = INT_MAX;
  }
  __syncthreads();
  for (int size = 2; size <= kPatCand; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int q = 0; q < kPatCand / 128; ++q) {
        const int p = lane + 64 * q;                        // pair index
        const int i = 2 * p - (p & (stride - 1));           // lower element of the pair
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
  if (r >= n) return;
  int32_t cnt = 0, ucnt = 0;
  const int32_t base = PASS == 1 ? rowptr[r] : 0;
  for (int q = 0; q < kPatCand / 64; ++q) {
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

// a numbering handed over by the caller: rep[row] = any (cell, local DOF) of
// the row (all writers store valid representatives), then every occurrence is
// checked to sit on the same vef at the same position as the representative
__global__ void k_rep_any(FaceTab t, int64_t total, const int32_t *__restrict__ dofs, int64_t *__restrict__ rep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int32_t d = dofs[i];
  if (d >= 0) rep[d] = i;
}

__device__ __forceinline__ void vef_key(const FaceTab &t, int64_t i, int (&key)[5]) {
  const int64_t cell = i / t.nd;
  const int l = (int)(i % t.nd), f = t.face_of[l];
  int ijk[3];
  cell_ijk(t, cell, ijk);
  key[0] = t.type[f];
  key[1] = t.pos_of[l];
  for (int k = 0; k < 3; ++k) key[2 + k] = ijk[k] + ((t.anchor[f] >> k) & 1);
}

__global__ void k_check_conforming(FaceTab t, int64_t total, const int32_t *__restrict__ dofs,
                                   const int64_t *__restrict__ rep, int32_t n, int32_t *__restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int32_t d = dofs[i];
  if (d < 0) return;
  if (d >= n || rep[d] < 0) {
    *bad = 1;
    return;
  }
  int a[5], b[5];
  vef_key(t, i, a);
  vef_key(t, rep[d], b);
  for (int k = 0; k < 5; ++k)
    if (a[k] != b[k]) *bad = 1;
}

__global__ void k_fill_i64(int64_t *__restrict__ p, int64_t n, int64_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------- symmetric layout test ---
// per block: the longest row (first such row), for the host to read D from
__global__ void k_longest_row(int32_t n, const int32_t *__restrict__ rowptr, int64_t *__restrict__ part) {
  __shared__ int64_t sm[256];
  const int32_t r = blockIdx.x * 256 + threadIdx.x;
  // key: length high, -row low, so the max is the longest, first row
  int64_t key = -1;
  if (r < n) key = ((int64_t)(rowptr[r + 1] - rowptr[r]) << 32) | (uint32_t)(INT32_MAX - r);
  sm[threadIdx.x] = key;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sm[threadIdx.x] = max(sm[threadIdx.x], sm[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

struct OffsetSet {
  int32_t k;
  int32_t o[64];
};

// every offset of every row is in D and the pattern is structurally symmetric
__global__ void k_check_offsets(int32_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                OffsetSet D, int32_t *__restrict__ bad) {
  const int32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  bool ok = true;
  for (int32_t e = rowptr[r]; e < rowptr[r + 1] && ok; ++e) {
    const int32_t c = col[e], o = c - r;
    int lo = 0, hi = D.k;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (D.o[mid] < o) lo = mid + 1;
      else hi = mid;
    }
    ok = lo < D.k && D.o[lo] == o;
    int32_t a = rowptr[c], b = rowptr[c + 1];   // (c, r) present
    while (a < b) {
      const int32_t mid = (a + b) >> 1;
      if (col[mid] < r) a = mid + 1;
      else b = mid;
    }
    ok = ok && a < rowptr[c + 1] && col[a] == r;
  }
  if (!ok) *bad = 1;
}

// ---------------------------------------------------- Dirichlet values ---
// value of the linear g at each fixed DOF's node, taken from the last cell
// that touches the node (the reference's cell loop leaves the last write);
// same expression as the node position of the FE space
// (sbm_serial_fe_space.i90:997-1176), without FMA contraction
__global__ void k_interp_linear(FaceTab t, int64_t ncells, const int32_t *__restrict__ dofs, double *__restrict__ out) {
#pragma clang fp contract(off)
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  int ijk[3];
  cell_ijk(t, c, ijk);
  for (int l = 0; l < t.nd; ++l) {
    const int32_t d = dofs[c * t.nd + l];
    if (d >= 0) continue;
    const int f = t.face_of[l];
    bool last = true;
    for (int k = 0; k < t.dim; ++k) {
      if ((t.type[f] >> k) & 1) continue;
      const int w = ijk[k] + ((t.anchor[f] >> k) & 1);
      last = last && ijk[k] == min(w, t.n[k] - 1);
    }
    if (!last) continue;
    int rem = l % t.nnode;
    double v = 0.0;
    for (int k = 0; k < t.dim; ++k) {
      const int m = rem % t.np1;
      rem /= t.np1;
      const double x0 = (t.hi[k] - t.lo[k]) * ((double)ijk[k] / (double)t.n[k]) + t.lo[k];
      const double x1 = (t.hi[k] - t.lo[k]) * ((double)(ijk[k] + 1) / (double)t.n[k]) + t.lo[k];
      const double s = (double)m / (double)t.order;
      v += x0 * (1.0 - s) + x1 * s;
    }
    out[-d - 1] = v;
  }
}

unsigned blocks(int64_t n, int b) { return (unsigned)std::max<int64_t>(1, (n + b - 1) / b); }

}  // namespace

void number_structured(femhip_space &s) {
  hipStream_t st = s.ctx->stream;
  const FaceTab t = face_tab(s);
  s.ncells = (int64_t)t.n[0] * t.n[1] * t.n[2];
  const int64_t nc = s.ncells;
  DevBuf<int32_t> nfree, nfixed;
  nfree.alloc((size_t)nc);
  nfixed.alloc((size_t)nc);
  hipLaunchKernelGGL(k_count_new, dim3(blocks(nc, 256)), dim3(256), 0, st, t, nc, nfree.p, nfixed.p);
  FH_HIP(hipGetLastError());
  const int64_t n_free = exclusive_scan(st, nfree.p, nfree.p, nc, false);
  const int64_t n_fixed = exclusive_scan(st, nfixed.p, nfixed.p, nc, false);
  s.d_dofs.alloc((size_t)(nc * t.nd));
  s.d_rep.alloc((size_t)std::max<int64_t>(1, n_free));
  hipLaunchKernelGGL(k_number, dim3(blocks(nc, 256)), dim3(256), 0, st, t, nc, nfree.p, nfixed.p, s.d_dofs.p,
                     s.d_rep.p);
  FH_HIP(hipGetLastError());
  FH_HIP(hipStreamSynchronize(st));
  s.n_free = (int32_t)n_free;
  s.n_fixed = (int32_t)n_fixed;
  s.dofs.clear();   // host copy made on request
}

void representatives_from_dofs(femhip_space &s) {
  hipStream_t st = s.ctx->stream;
  const FaceTab t = face_tab(s);
  const int64_t total = s.ncells * t.nd;
  s.d_rep.alloc((size_t)std::max(1, s.n_free));
  hipLaunchKernelGGL(k_fill_i64, dim3(blocks(s.n_free, 256)), dim3(256), 0, st, s.d_rep.p, (int64_t)s.n_free,
                     (int64_t)-1);
  hipLaunchKernelGGL(k_rep_any, dim3(blocks(total, 256)), dim3(256), 0, st, t, total, s.d_dofs.p, s.d_rep.p);
  DevBuf<int32_t> bad;
  bad.alloc(1);
  FH_HIP(hipMemsetAsync(bad.p, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_check_conforming, dim3(blocks(total, 256)), dim3(256), 0, st, t, total, s.d_dofs.p, s.d_rep.p,
                     s.n_free, bad.p);
  FH_HIP(hipGetLastError());
  int32_t h_bad = 1;
  FH_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  FH_CHECK(!h_bad, FEMHIP_ERR_INVALID,
           "lst_dofs_gids does not conform to the mesh (a free id is missing or shared by different vefs)");
}

void build_pattern(femhip_operator &op) {
  femhip_space &s = *op.space;
  hipStream_t st = op.ctx->stream;
  const FaceTab t = face_tab(s);
  FH_CHECK((1 << t.dim) * t.nd <= kPatCand, FEMHIP_ERR_INVALID, "too many local DOFs for the pattern kernel");
  const int32_t n = s.n_free;
  op.n = n;
  op.rowptr.alloc((size_t)n + 1);
  DevBuf<int32_t> ulen;
  ulen.alloc((size_t)std::max(1, n));
  const unsigned nb = blocks(n, kPatWaves);
  hipLaunchKernelGGL(k_pattern<0>, dim3(nb), dim3(kPatWaves * 64), 0, st, t, n, s.d_rep.p, s.d_dofs.p, op.rowptr.p,
                     ulen.p, (const int32_t *)nullptr, (int32_t *)nullptr);
  FH_HIP(hipGetLastError());
  op.nnz_full = exclusive_scan(st, op.rowptr.p, op.rowptr.p, n, true);
  op.nnz_upper = exclusive_scan(st, ulen.p, ulen.p, n, false);
  op.col.alloc((size_t)op.nnz_full);
  hipLaunchKernelGGL(k_pattern<1>, dim3(nb), dim3(kPatWaves * 64), 0, st, t, n, s.d_rep.p, s.d_dofs.p,
                     (int32_t *)nullptr, (int32_t *)nullptr, op.rowptr.p, op.col.p);
  FH_HIP(hipGetLastError());
  FH_HIP(hipStreamSynchronize(st));
  op.h_rowptr.clear();
  op.h_col.clear();
}

void ensure_host_pattern(const femhip_operator &op) {
  if (!op.h_rowptr.empty() || op.structured) return;
  auto &o = const_cast<femhip_operator &>(op);
  FH_HIP(hipStreamSynchronize(op.ctx->stream));
  o.h_rowptr.resize((size_t)op.n + 1);
  o.h_col.resize((size_t)op.nnz_full);
  FH_HIP(hipMemcpy(o.h_rowptr.data(), op.rowptr.p, sizeof(int32_t) * ((size_t)op.n + 1), hipMemcpyDeviceToHost));
  FH_HIP(hipMemcpy(o.h_col.data(), op.col.p, sizeof(int32_t) * (size_t)op.nnz_full, hipMemcpyDeviceToHost));
}

void ensure_host_dofs(const femhip_space &s) {
  if (!s.dofs.empty()) return;
  auto &m = const_cast<femhip_space &>(s);
  const size_t cnt = (size_t)(s.ncells * s.fe.nd);
  FH_HIP(hipStreamSynchronize(s.ctx->stream));
  m.dofs.resize(cnt);
  FH_HIP(hipMemcpy(m.dofs.data(), s.d_dofs.p, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost));
  for (auto &g : m.dofs)
    if (g >= 0) ++g;   // device: free -> 0-based row; API: 1-based ids
}

bool symmetric_offsets(const femhip_operator &op, std::vector<int32_t> &D) {
  const int32_t n = op.n;
  if (n <= 0) return false;
  hipStream_t st = op.ctx->stream;
  const unsigned nb = blocks(n, 256);
  DevBuf<int64_t> part;
  part.alloc(nb);
  hipLaunchKernelGGL(k_longest_row, dim3(nb), dim3(256), 0, st, n, op.rowptr.p, part.p);
  FH_HIP(hipGetLastError());
  std::vector<int64_t> hp(nb);
  FH_HIP(hipMemcpyAsync(hp.data(), part.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  const int64_t best = *std::max_element(hp.begin(), hp.end());
  const int32_t r = INT32_MAX - (int32_t)(best & 0xffffffff), len = (int32_t)(best >> 32);
  if (len <= 0 || len > 64) return false;
  int32_t b0 = 0;
  D.resize((size_t)len);
  FH_HIP(hipMemcpy(&b0, op.rowptr.p + r, sizeof(int32_t), hipMemcpyDeviceToHost));
  FH_HIP(hipMemcpy(D.data(), op.col.p + b0, sizeof(int32_t) * (size_t)len, hipMemcpyDeviceToHost));
  for (auto &c : D) c -= r;   // columns ascending -> offsets ascending
  OffsetSet os{};
  os.k = len;
  for (int i = 0; i < len; ++i) os.o[i] = D[(size_t)i];
  DevBuf<int32_t> bad;
  bad.alloc(1);
  FH_HIP(hipMemsetAsync(bad.p, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_check_offsets, dim3(nb), dim3(256), 0, st, n, op.rowptr.p, op.col.p, os, bad.p);
  FH_HIP(hipGetLastError());
  int32_t h_bad = 1;
  FH_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  return h_bad == 0;
}

void interpolate_dirichlet_linear(const femhip_space &s, double *out_host) {
  if (s.n_fixed == 0) return;
  hipStream_t st = s.ctx->stream;
  const FaceTab t = face_tab(s);
  DevBuf<double> out;
  out.alloc((size_t)s.n_fixed);
  hipLaunchKernelGGL(k_interp_linear, dim3(blocks(s.ncells, 256)), dim3(256), 0, st, t, s.ncells, s.d_dofs.p, out.p);
  FH_HIP(hipGetLastError());
  FH_HIP(hipMemcpyAsync(out_host, out.p, sizeof(double) * (size_t)s.n_fixed, hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
}

}  // namespace femhip
