// sliced.cpp -- host construction of the sliced diagonal/ELL solver format from
// the CSR pattern (pattern only; values are gathered on the device through
// `map` after every assembly).
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdlib>
#include <thread>

#include "femhip_internal.h"

namespace femhip {

namespace {

template <typename F>
void parallel_rows(int64_t n, F &&f) {
  const unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nth; ++t)
    th.emplace_back([&, t] { f(t, n * t / nth, n * (t + 1) / nth); });
  for (auto &x : th) x.join();
}

// Symmetric-diagonal layout if the pattern is structurally symmetric and all
// rows draw their offsets from one small symmetric set (checked on the device
// against the longest row's offsets).
bool try_build_symmetric_dia(femhip_operator &op) {
  Sliced &S = op.sl;
  const int32_t n = op.n;
  if (n == 0) return false;
  std::vector<int32_t> D;
  if (!symmetric_offsets(op, D)) return false;
  for (int32_t o : D)
    if (!std::binary_search(D.begin(), D.end(), -o)) return false;
  std::vector<int32_t> up;
  for (int32_t o : D)
    if (o >= 0) up.push_back(o);
  if (up.empty() || up[0] != 0) return false;
  const int64_t stride = ((int64_t)n + kSlice - 1) / kSlice * kSlice;
  const int64_t E = (int64_t)up.size() * stride;
  const int64_t stored = (op.nnz_full + n) / 2;
  if (E > stored * 3 / 2 + 64 || E >= (int64_t)INT32_MAX) return false;
  // values are filled from the CSR row by row on the device (no entry map)
  S.sym = true;
  S.kplus = (int32_t)up.size();
  S.stride = stride;
  S.h_soffs = up;
  S.ns = (int32_t)(stride / kSlice);
  S.entries = E;
  S.dia_slices = S.ns;
  S.soffs.alloc(up.size());
  FH_HIP(hipMemcpy(S.soffs.p, up.data(), sizeof(int32_t) * up.size(), hipMemcpyHostToDevice));
  // unused sliced arrays stay valid 1-element buffers
  S.map.alloc(1);
  S.base.alloc(1);
  S.meta.alloc(1);
  S.aux.alloc(1);
  S.offs.alloc(1);
  S.ecol.alloc(1);
  S.built = true;
  S.rows = n;
  S.permuted = false;
  S.filled_version = S.filled_version32 = -1;
  return true;
}

// Slices of the rows (rp, cl) (n rows, columns sorted); value entries map to
// CSR position pos[k] (pos null: k itself; -2 in pos: the value 1).
void build_slices(femhip_operator &op, int32_t n, const std::vector<int32_t> &rp, const std::vector<int32_t> &cl,
                  const std::vector<int32_t> *pos);

constexpr int kClassMaxOffsets = 128;

// every row r draws its column offsets perm[c] - perm[r] from its class's set
__global__ void k_class_check(int32_t n, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                              const int32_t *__restrict__ perm, const int32_t *__restrict__ cls,
                              const int32_t *__restrict__ offs, const int32_t *__restrict__ optr,
                              int32_t *__restrict__ bad) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int c = cls[r];
  const int32_t o0 = optr[c], o1 = optr[c + 1], pr = perm[r];
  bool ok = true;
  for (int32_t e = rowptr[r]; e < rowptr[r + 1] && ok; ++e) {
    const int32_t o = perm[col[e]] - pr;
    int32_t lo = o0, hi = o1;
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (offs[mid] < o) lo = mid + 1;
      else hi = mid;
    }
    ok = lo < o1 && offs[lo] == o;
  }
  if (!ok) *bad = 1;
}

// per slice (64 solver rows, one class): which of the class's offsets its rows
// use, as a bit mask over the class's sorted offsets (padding rows use 0)
__global__ __launch_bounds__(kSlice) void k_class_mask(int32_t rows, const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ col,
                                                       const int32_t *__restrict__ perm,
                                                       const int32_t *__restrict__ iperm,
                                                       const int32_t *__restrict__ slice_cls,
                                                       const int32_t *__restrict__ offs,
                                                       const int32_t *__restrict__ optr, uint64_t *__restrict__ mask) {
  __shared__ uint64_t sm[2][kSlice];
  const int32_t sl = blockIdx.x, lane = threadIdx.x, i = sl * kSlice + lane;
  const int c = slice_cls[sl];
  const int32_t o0 = optr[c], o1 = optr[c + 1];
  uint64_t m[2] = {0, 0};
  auto set = [&](int32_t o) {
    int32_t lo = o0, hi = o1;
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (offs[mid] < o) lo = mid + 1;
      else hi = mid;
    }
    if (lo < o1 && offs[lo] == o) m[(lo - o0) >> 6] |= 1ull << ((lo - o0) & 63);
  };
  const int32_t r = i < rows ? iperm[i] : -1;
  if (r >= 0) {
    for (int32_t e = rowptr[r]; e < rowptr[r + 1]; ++e) set(perm[col[e]] - i);
  } else if (i < rows) {
    set(0);
  }
  sm[0][lane] = m[0];
  sm[1][lane] = m[1];
  __syncthreads();
  if (lane == 0) {
    uint64_t a = 0, b = 0;
    for (int t = 0; t < kSlice; ++t) {
      a |= sm[0][t];
      b |= sm[1][t];
    }
    mask[2 * (int64_t)sl] = a;
    mask[2 * (int64_t)sl + 1] = b;
  }
}

// map of the class-blocked slices: entry (slice, k, lane) of solver row i holds
// the CSR position of (iperm[i], iperm[i + offset k]), -2 for the unit
// diagonal of a padding row, -1 where there is no entry
__global__ void k_class_map(int32_t rows, const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                            const int32_t *__restrict__ iperm, const int64_t *__restrict__ base,
                            const int32_t *__restrict__ meta, const int32_t *__restrict__ aux,
                            const int32_t *__restrict__ offs, int32_t *__restrict__ map) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const int32_t sl = i / kSlice, lane = i % kSlice;
  const int32_t K = meta[sl] & 0x0fffffff, a = aux[sl];
  const int32_t r = iperm[i];
  const int32_t b = r >= 0 ? rowptr[r] : 0, e = r >= 0 ? rowptr[r + 1] : 0;
  for (int32_t k = 0; k < K; ++k) {
    const int32_t o = offs[a + k], j = i + o;
    int32_t m = -1;
    if (r < 0) {
      m = o == 0 ? -2 : -1;
    } else if (j >= 0 && j < rows && iperm[j] >= 0) {
      const int32_t c = iperm[j];
      int32_t lo = b, hi = e;
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (col[mid] < c) lo = mid + 1;
        else hi = mid;
      }
      if (lo < e && col[lo] == c) m = lo;
    }
    map[base[sl] + (int64_t)k * kSlice + lane] = m;
  }
}

// Class-blocked diagonal layout of a structured space of order p >= 2, or of a
// vector space of any order.  The free DOFs sit on the nodes of the fine grid
// (p n_k + 1 nodes along k); a DOF's class is its component and its node's
// coordinate modulo p in every direction (p^d classes per component:
// vertices, edge, face and cell nodes for Q2; one per component for vector
// Q1) and its class coordinate the node coordinate divided by p.  The solver numbers the rows class by class, each
// class lexicographically in a box whose x (and y) strides are the same for
// all classes, so the column offsets from a row of class A to its neighbours
// of class B do not depend on the row: every 64-row slice of one class has
// one set of diagonals and needs no column indices (64^3 Q2: 1.11 GB a
// product instead of 1.59 GB with the first-touch rows in ELL slices).
// Box positions without a free DOF become identity rows.  The operator's
// numbering, CSR and values are untouched: the solver permutes b and x.  The
// permutation is computed on the host from the rows' representatives; the
// offsets check and the slices' value map are built on the device.
bool try_build_class_dia(femhip_operator &op) {
  const char *env = getenv("FEMHIP_CLASS_DIA");   // read per operator: tests compare both layouts
  if (env && atoi(env) == 0) return false;
  if (!op.space || op.structured || op.n <= 0) return false;
  femhip_space &sp = *op.space;
  const RefFE &fe = sp.fe;
  if (fe.order < 2 && fe.ncomp == 1) return false;   // scalar Q1: the symmetric-diagonal layout
  const int dim = fe.dim, p = fe.order, nd = fe.nd;
  const int32_t n = op.n;
  int npar = 1;   // node-parity classes per component
  for (int k = 0; k < dim; ++k) npar *= p;
  const int ncls = npar * fe.ncomp;
  std::vector<int64_t> rep((size_t)n);
  FH_HIP(hipMemcpy(rep.data(), sp.d_rep.p, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToHost));
  const int mn[3] = {sp.mesh.n[0], dim > 1 ? sp.mesh.n[1] : 1, dim > 2 ? sp.mesh.n[2] : 1};
  std::vector<int32_t> cls((size_t)n), g((size_t)n * 3, 0);
  std::vector<int32_t> lo((size_t)ncls * 3, INT32_MAX), hi((size_t)ncls * 3, -1);
  for (int32_t r = 0; r < n; ++r) {
    const int64_t cell = rep[(size_t)r] / nd;
    const int l = (int)(rep[(size_t)r] % nd), node = l % fe.nnode;   // local DOF = component * nodes + node
    int64_t cr = cell;
    int c = (l / fe.nnode) * npar, w = 1, nr = node;
    for (int k = 0; k < dim; ++k) {
      const int ijk = (int)(cr % mn[k]);
      cr /= mn[k];
      const int a = p * ijk + nr % fe.np1;   // fine-grid node coordinate
      nr /= fe.np1;
      c += (a % p) * w;
      w *= p;
      g[(size_t)r * 3 + k] = a / p;
    }
    cls[(size_t)r] = c;
    for (int k = 0; k < dim; ++k) {
      lo[(size_t)c * 3 + k] = std::min(lo[(size_t)c * 3 + k], g[(size_t)r * 3 + k]);
      hi[(size_t)c * 3 + k] = std::max(hi[(size_t)c * 3 + k], g[(size_t)r * 3 + k]);
    }
  }
  int64_t stride[3] = {1, 1, 1};   // common box extents of all but the last direction
  for (int k = 0; k + 1 < dim; ++k)
    for (int c = 0; c < ncls; ++c)
      if (hi[(size_t)c * 3 + k] >= 0) stride[k] = std::max<int64_t>(stride[k], hi[(size_t)c * 3 + k] - lo[(size_t)c * 3 + k] + 1);
  std::vector<int64_t> base((size_t)ncls + 1, 0);
  for (int c = 0; c < ncls; ++c) {
    int64_t size = 0;
    if (hi[(size_t)c * 3 + dim - 1] >= 0) {
      size = hi[(size_t)c * 3 + dim - 1] - lo[(size_t)c * 3 + dim - 1] + 1;
      for (int k = 0; k + 1 < dim; ++k) size *= stride[k];
    }
    base[(size_t)c + 1] = base[(size_t)c] + (size + kSlice - 1) / kSlice * kSlice;   // classes start on a slice
  }
  const int64_t rows = base[(size_t)ncls];
  if (rows >= (int64_t)INT32_MAX || rows > 2 * (int64_t)n + 64 * ncls) return false;
  std::vector<int32_t> perm((size_t)n), iperm((size_t)rows, -1);
  for (int32_t r = 0; r < n; ++r) {
    const int c = cls[(size_t)r];
    int64_t idx = 0, mul = 1;
    for (int k = 0; k < dim; ++k) {
      idx += mul * (g[(size_t)r * 3 + k] - lo[(size_t)c * 3 + k]);
      mul *= stride[k];
    }
    perm[(size_t)r] = (int32_t)(base[(size_t)c] + idx);
    FH_CHECK(iperm[(size_t)perm[(size_t)r]] < 0, FEMHIP_ERR_INVALID, "two free DOFs on one node");
    iperm[(size_t)perm[(size_t)r]] = r;
  }
  // each class's offsets, read off its longest row; every row of the class is
  // checked on the device to draw its offsets from that set
  std::vector<int32_t> rowlen((size_t)n + 1);
  FH_HIP(hipMemcpy(rowlen.data(), op.rowptr.p, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost));
  std::vector<int32_t> best((size_t)ncls, -1);
  for (int32_t r = 0; r < n; ++r) {
    int32_t &b = best[(size_t)cls[(size_t)r]];
    if (b < 0 || rowlen[(size_t)r + 1] - rowlen[(size_t)r] > rowlen[(size_t)b + 1] - rowlen[(size_t)b]) b = r;
  }
  std::vector<int32_t> offs, off_ptr((size_t)ncls + 1, 0);
  for (int c = 0; c < ncls; ++c) {
    const int32_t r = best[(size_t)c];
    if (r >= 0) {
      const int32_t len = rowlen[(size_t)r + 1] - rowlen[(size_t)r];
      if (len > kClassMaxOffsets) return false;
      std::vector<int32_t> cols((size_t)len);
      FH_HIP(hipMemcpy(cols.data(), op.col.p + rowlen[(size_t)r], sizeof(int32_t) * (size_t)len,
                       hipMemcpyDeviceToHost));
      std::vector<int32_t> o;
      for (int32_t col : cols) o.push_back(perm[(size_t)col] - perm[(size_t)r]);
      std::sort(o.begin(), o.end());
      offs.insert(offs.end(), o.begin(), o.end());
    }
    off_ptr[(size_t)c + 1] = (int32_t)offs.size();
  }
  rowlen.clear();
  rowlen.shrink_to_fit();
  Sliced &S = op.sl;
  hipStream_t st = op.ctx->stream;
  DevBuf<int32_t> d_cls, d_iperm, d_offs, d_optr, bad;
  S.perm.alloc((size_t)n);
  d_cls.alloc((size_t)n);
  d_iperm.alloc((size_t)rows);
  d_offs.alloc(std::max<size_t>(1, offs.size()));
  d_optr.alloc((size_t)ncls + 1);
  bad.alloc(1);
  FH_HIP(hipMemcpy(S.perm.p, perm.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
  FH_HIP(hipMemcpy(d_cls.p, cls.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
  FH_HIP(hipMemcpy(d_iperm.p, iperm.data(), sizeof(int32_t) * (size_t)rows, hipMemcpyHostToDevice));
  if (!offs.empty()) FH_HIP(hipMemcpy(d_offs.p, offs.data(), sizeof(int32_t) * offs.size(), hipMemcpyHostToDevice));
  FH_HIP(hipMemcpy(d_optr.p, off_ptr.data(), sizeof(int32_t) * off_ptr.size(), hipMemcpyHostToDevice));
  FH_HIP(hipMemsetAsync(bad.p, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_class_check, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, op.rowptr.p, op.col.p,
                     S.perm.p, d_cls.p, d_offs.p, d_optr.p, bad.p);
  FH_HIP(hipGetLastError());
  int32_t h_bad = 1;
  FH_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  FH_HIP(hipStreamSynchronize(st));
  if (h_bad) {
    S.perm.release();
    return false;
  }
  // slices: every slice lies in one class and keeps the class offsets its
  // rows use (diagonal slices, no column indices: boundary slices take fewer
  // offsets); padding rows keep their unit diagonal
  const int32_t ns = (int32_t)(rows / kSlice);
  std::vector<int32_t> scls((size_t)ns);
  for (int32_t sl = 0, c = 0; sl < ns; ++sl) {
    while (base[(size_t)c + 1] <= (int64_t)sl * kSlice) ++c;
    scls[(size_t)sl] = c;
  }
  std::vector<uint64_t> hmask((size_t)ns * 2);
  {
    DevBuf<int32_t> d_scls;
    DevBuf<uint64_t> d_mask;
    d_scls.alloc((size_t)ns);
    d_mask.alloc((size_t)ns * 2);
    FH_HIP(hipMemcpy(d_scls.p, scls.data(), sizeof(int32_t) * (size_t)ns, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_class_mask, dim3((unsigned)ns), dim3(kSlice), 0, st, (int32_t)rows, op.rowptr.p, op.col.p,
                       S.perm.p, d_iperm.p, d_scls.p, d_offs.p, d_optr.p, d_mask.p);
    FH_HIP(hipGetLastError());
    FH_HIP(hipMemcpyAsync(hmask.data(), d_mask.p, sizeof(uint64_t) * hmask.size(), hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
  }
  std::vector<int64_t> sbase((size_t)ns + 1, 0);
  std::vector<int32_t> meta((size_t)ns), aux((size_t)ns), soffs;
  for (int32_t sl = 0; sl < ns; ++sl) {
    const int c = scls[(size_t)sl];
    aux[(size_t)sl] = (int32_t)soffs.size();
    for (int32_t t = off_ptr[(size_t)c]; t < off_ptr[(size_t)c + 1]; ++t) {
      const int bit = t - off_ptr[(size_t)c];
      if ((hmask[(size_t)sl * 2 + (bit >> 6)] >> (bit & 63)) & 1) soffs.push_back(offs[(size_t)t]);
    }
    const int32_t K = (int32_t)soffs.size() - aux[(size_t)sl];
    sbase[(size_t)sl + 1] = sbase[(size_t)sl] + (int64_t)K * kSlice;
    meta[(size_t)sl] = K;
    if (K > 0) {   // bit 29: some lane's column row + offset leaves [0, rows)
      const int64_t lo = (int64_t)sl * kSlice + soffs[(size_t)aux[(size_t)sl]];
      const int64_t hi = (int64_t)sl * kSlice + kSlice - 1 + soffs.back();
      if (lo < 0 || hi >= rows) meta[(size_t)sl] |= 1 << 29;
    }
  }
  offs.swap(soffs);
  const int64_t E = sbase[(size_t)ns];
  if (E >= (int64_t)INT32_MAX) {
    S.perm.release();
    return false;
  }
  S.sym = false;
  S.perm_slices = 0;
  S.ns = ns;
  S.entries = E;
  S.dia_slices = ns;
  S.base.alloc((size_t)ns + 1);
  S.meta.alloc((size_t)std::max(1, ns));
  S.aux.alloc((size_t)std::max(1, ns));
  S.offs.alloc(std::max<size_t>(1, offs.size()));
  S.ecol.alloc(1);
  S.srow.alloc(1);
  S.map.alloc((size_t)std::max<int64_t>(1, E));
  FH_HIP(hipMemcpy(S.base.p, sbase.data(), sizeof(int64_t) * sbase.size(), hipMemcpyHostToDevice));
  FH_HIP(hipMemcpy(S.meta.p, meta.data(), sizeof(int32_t) * meta.size(), hipMemcpyHostToDevice));
  FH_HIP(hipMemcpy(S.aux.p, aux.data(), sizeof(int32_t) * aux.size(), hipMemcpyHostToDevice));
  if (!offs.empty()) FH_HIP(hipMemcpy(S.offs.p, offs.data(), sizeof(int32_t) * offs.size(), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_class_map, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st, (int32_t)rows,
                     op.rowptr.p, op.col.p, d_iperm.p, S.base.p, S.meta.p, S.aux.p, S.offs.p, S.map.p);
  FH_HIP(hipGetLastError());
  FH_HIP(hipStreamSynchronize(st));
  S.built = true;
  S.rows = (int32_t)rows;
  S.permuted = true;
  S.filled_version = S.filled_version32 = -1;
  return true;
}

}  // namespace

void build_sliced(femhip_operator &op) {
  // symmetric-diagonal storage wherever the pattern allows it (half the matrix
  // stream); FEMHIP_SYMDIA=0 keeps the full diagonals
  static const char *sym_env = getenv("FEMHIP_SYMDIA");
  const bool want_sym = sym_env ? atoi(sym_env) != 0 : true;
  if (want_sym && try_build_symmetric_dia(op)) return;
  if (try_build_class_dia(op)) return;
  ensure_host_pattern(op);
  build_slices(op, op.n, op.h_rowptr, op.h_col, nullptr);
}

namespace {

void build_slices(femhip_operator &op, int32_t n, const std::vector<int32_t> &rp, const std::vector<int32_t> &cl,
                  const std::vector<int32_t> *pos) {
  Sliced &S = op.sl;
  S.sym = false;
  S.permuted = false;
  S.perm_slices = 0;
  const int32_t ns = (n + kSlice - 1) / kSlice;
  auto at = [&](int32_t k) { return pos ? (*pos)[(size_t)k] : k; };

  // per slice: distinct offsets, longest row, storage choice
  std::vector<std::vector<int32_t>> offsets((size_t)ns);
  std::vector<int32_t> K((size_t)ns), is_ell((size_t)ns);
  unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  auto work = [&](unsigned t) {
    std::vector<int32_t> buf;
    for (int32_t s = (int32_t)((int64_t)ns * t / nth); s < (int32_t)((int64_t)ns * (t + 1) / nth); ++s) {
      buf.clear();
      int32_t maxlen = 0;
      for (int32_t r = s * kSlice; r < std::min(n, (s + 1) * kSlice); ++r) {
        maxlen = std::max(maxlen, rp[(size_t)r + 1] - rp[(size_t)r]);
        for (int32_t k = rp[(size_t)r]; k < rp[(size_t)r + 1]; ++k) buf.push_back(cl[(size_t)k] - r);
      }
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      // diagonals pay off when they cost at most ~25% more values than ELL
      // (they save the 4-byte column index per entry)
      const bool dia = (int64_t)buf.size() * 4 <= (int64_t)maxlen * 5 + 4;
      is_ell[(size_t)s] = dia ? 0 : 1;
      K[(size_t)s] = dia ? (int32_t)buf.size() : maxlen;
      if (dia) offsets[(size_t)s] = buf;
    }
  };
  {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; ++t) th.emplace_back(work, t);
    for (auto &x : th) x.join();
  }
  // a window's rows are sorted by length (longest first, ties by row) and cut
  // into ELL slices when that stores fewer bytes than the slices as they are
  // (values at 8 B, ELL columns and row lists at 4 B)
  static const char *win_env = getenv("FEMHIP_SORT_WINDOW");
  const int32_t window = win_env ? std::max(0, atoi(win_env)) / kSlice * kSlice : kSortWindow;
  const int32_t wsl = std::max(1, window / kSlice);   // slices per window
  std::vector<int32_t> rowof((size_t)ns * kSlice);
  std::vector<char> perm((size_t)ns, 0);
  bool any_perm = false;
  for (int32_t w0 = 0; w0 < ns; w0 += wsl) {
    const int32_t w1 = std::min(ns, w0 + wsl);
    const int32_t r0 = w0 * kSlice, r1 = std::min(n, w1 * kSlice);
    for (int32_t i = r0; i < w1 * kSlice; ++i) rowof[(size_t)i] = i < r1 ? i : n;
    if (window == 0) continue;
    int64_t as_is = 0;
    for (int32_t s = w0; s < w1; ++s) as_is += (int64_t)K[(size_t)s] * kSlice * (is_ell[(size_t)s] ? 12 : 8);
    std::vector<int32_t> sorted(rowof.begin() + r0, rowof.begin() + r1);
    std::stable_sort(sorted.begin(), sorted.end(), [&](int32_t a, int32_t b) {
      return rp[(size_t)a + 1] - rp[(size_t)a] > rp[(size_t)b + 1] - rp[(size_t)b];
    });
    int64_t if_sorted = 0;
    for (size_t i = 0; i < sorted.size(); i += kSlice)
      if_sorted += (int64_t)(rp[(size_t)sorted[i] + 1] - rp[(size_t)sorted[i]]) * kSlice * 12 + 4 * kSlice;
    if (if_sorted >= as_is) continue;
    any_perm = true;
    std::copy(sorted.begin(), sorted.end(), rowof.begin() + r0);
    for (int32_t s = w0; s < w1; ++s) {
      const int32_t first = rowof[(size_t)s * kSlice];
      is_ell[(size_t)s] = 1;
      perm[(size_t)s] = 1;
      ++S.perm_slices;
      K[(size_t)s] = first < n ? rp[(size_t)first + 1] - rp[(size_t)first] : 0;   // longest row of the slice
      offsets[(size_t)s].clear();
    }
  }
  std::vector<int64_t> base((size_t)ns + 1, 0);
  std::vector<int32_t> meta((size_t)ns), aux((size_t)ns);
  std::vector<int32_t> offs;
  int64_t ell_entries = 0, dia = 0;
  for (int32_t s = 0; s < ns; ++s) {
    base[(size_t)s + 1] = base[(size_t)s] + (int64_t)K[(size_t)s] * kSlice;
    FH_CHECK(K[(size_t)s] < (1 << 28), FEMHIP_ERR_INVALID, "row too long for the sliced format");
    meta[(size_t)s] = K[(size_t)s] | (is_ell[(size_t)s] << 30) | (perm[(size_t)s] << 28);
    if (!is_ell[(size_t)s] && !offsets[(size_t)s].empty()) {
      // bit 29: some lane's column row + offset leaves [0, n) (padding near the ends)
      const int64_t lo = (int64_t)s * kSlice + offsets[(size_t)s].front();
      const int64_t hi = (int64_t)s * kSlice + kSlice - 1 + offsets[(size_t)s].back();
      if (lo < 0 || hi >= n) meta[(size_t)s] |= 1 << 29;
    }
    if (is_ell[(size_t)s]) {
      aux[(size_t)s] = (int32_t)ell_entries;
      ell_entries += (int64_t)K[(size_t)s] * kSlice;
    } else {
      aux[(size_t)s] = (int32_t)offs.size();
      offs.insert(offs.end(), offsets[(size_t)s].begin(), offsets[(size_t)s].end());
      ++dia;
    }
  }
  const int64_t E = base[(size_t)ns];
  FH_CHECK(E < (int64_t)INT32_MAX && ell_entries < (int64_t)INT32_MAX, FEMHIP_ERR_INVALID,
           "sliced matrix exceeds 2^31 entries");
  std::vector<int32_t> map((size_t)E, -1), ecol((size_t)std::max<int64_t>(1, ell_entries), 0);
  auto fill = [&](unsigned t) {
    for (int32_t s = (int32_t)((int64_t)ns * t / nth); s < (int32_t)((int64_t)ns * (t + 1) / nth); ++s) {
      const int32_t k_s = K[(size_t)s];
      for (int32_t lane = 0; lane < kSlice; ++lane) {
        const int32_t r = rowof[(size_t)s * kSlice + lane];
        if (r >= n) {
          if (is_ell[(size_t)s])
            for (int32_t k = 0; k < k_s; ++k) ecol[(size_t)(aux[(size_t)s] + (int64_t)k * kSlice + lane)] = n - 1;
          continue;
        }
        const int32_t b = rp[(size_t)r], e = rp[(size_t)r + 1];
        if (is_ell[(size_t)s]) {
          for (int32_t k = 0; k < k_s; ++k) {
            const int64_t idx = base[(size_t)s] + (int64_t)k * kSlice + lane;
            const int64_t ci = aux[(size_t)s] + (int64_t)k * kSlice + lane;
            if (b + k < e) {
              map[(size_t)idx] = at(b + k);
              ecol[(size_t)ci] = cl[(size_t)(b + k)];
            } else {
              ecol[(size_t)ci] = r;   // padding: valid column, zero value
            }
          }
        } else {
          const std::vector<int32_t> &off = offsets[(size_t)s];
          int32_t p = b;
          for (int32_t k = 0; k < k_s; ++k) {   // both sorted: merge
            const int32_t c = r + off[(size_t)k];
            while (p < e && cl[(size_t)p] < c) ++p;
            if (p < e && cl[(size_t)p] == c) map[(size_t)(base[(size_t)s] + (int64_t)k * kSlice + lane)] = at(p);
          }
        }
      }
    }
  };
  {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; ++t) th.emplace_back(fill, t);
    for (auto &x : th) x.join();
  }
  S.ns = ns;
  S.entries = E;
  S.dia_slices = dia;
  S.base.alloc((size_t)ns + 1);
  S.meta.alloc((size_t)std::max(1, ns));
  S.aux.alloc((size_t)std::max(1, ns));
  S.offs.alloc(std::max<size_t>(1, offs.size()));
  S.ecol.alloc(ecol.size());
  S.map.alloc((size_t)std::max<int64_t>(1, E));
  S.srow.alloc(any_perm ? rowof.size() : 1);
  if (any_perm) FH_HIP(hipMemcpy(S.srow.p, rowof.data(), sizeof(int32_t) * rowof.size(), hipMemcpyHostToDevice));
  FH_HIP(hipMemcpy(S.base.p, base.data(), sizeof(int64_t) * base.size(), hipMemcpyHostToDevice));
  if (ns) {
    FH_HIP(hipMemcpy(S.meta.p, meta.data(), sizeof(int32_t) * meta.size(), hipMemcpyHostToDevice));
    FH_HIP(hipMemcpy(S.aux.p, aux.data(), sizeof(int32_t) * aux.size(), hipMemcpyHostToDevice));
  }
  if (!offs.empty()) FH_HIP(hipMemcpy(S.offs.p, offs.data(), sizeof(int32_t) * offs.size(), hipMemcpyHostToDevice));
  FH_HIP(hipMemcpy(S.ecol.p, ecol.data(), sizeof(int32_t) * ecol.size(), hipMemcpyHostToDevice));
  if (E) FH_HIP(hipMemcpy(S.map.p, map.data(), sizeof(int32_t) * (size_t)E, hipMemcpyHostToDevice));
  S.built = true;
  S.rows = n;
  S.filled_version = S.filled_version32 = -1;
}

}  // namespace

}  // namespace femhip
