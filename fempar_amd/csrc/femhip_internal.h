// femhip_internal.h -- shared internal declarations of libfemhip (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "femhip.h"

namespace femhip {

// ---------------------------------------------------------------- errors ---
void set_error(const std::string &msg);
struct Error {
  int code;
  std::string msg;
};

#define FH_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw ::femhip::Error{FEMHIP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define FH_CHECK(cond, code, msg)                                                     \
  do {                                                                                \
    if (!(cond)) throw ::femhip::Error{(code), (msg)};                                \
  } while (0)

// ------------------------------------------------------- device buffers ---
// A block of device memory handed out by a bump pointer.  While a thread has
// one installed (ArenaScope), DevBuf::alloc carves its buffers from it
// (256-byte aligned) and release gives nothing back: the block's owner reuses
// it as a whole.  An allocation that does not fit falls back to hipMalloc and
// is counted.  The request server installs each request's block while it sets
// the request up, so no hipMalloc or hipFree runs while other requests'
// kernels are launched.
struct DevArena {
  char *base = nullptr;
  size_t size = 0, used = 0;
  size_t misses = 0;
};
DevArena *&current_arena();   // per thread
struct ArenaScope {
  DevArena *prev;
  explicit ArenaScope(DevArena *a) : prev(current_arena()) { current_arena() = a; }
  ~ArenaScope() { current_arena() = prev; }
  ArenaScope(const ArenaScope &) = delete;
  ArenaScope &operator=(const ArenaScope &) = delete;
};

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  bool owned = true;   // false: carved from an arena
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    const size_t bytes = count * sizeof(T);
    if (DevArena *a = current_arena()) {
      const size_t off = (a->used + 255) & ~(size_t)255;
      if (off + bytes <= a->size) {
        p = reinterpret_cast<T *>(a->base + off);
        a->used = off + bytes;
        n = count;
        owned = false;
        return;
      }
      ++a->misses;
    }
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error{FEMHIP_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed"};
    }
    n = count;
    owned = true;
  }
  void release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    n = 0;
    owned = true;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// ------------------------------------------------------------ reference FE ---
constexpr int kMaxNd = 81;   // Q2 vector 3D upper bound of local DOFs
constexpr int kMaxNq = 27;   // 3x3x3 Gauss points
constexpr int kMaxNf = 27;   // n-faces of a hexahedron incl. the cell

struct RefFE {
  int dim = 0, order = 0, ncomp = 1;
  int np1 = 0, nnode = 0, nd = 0, nq1 = 0, nq = 0, nnf = 0;
  int nf_type[kMaxNf], nf_anchor[kMaxNf];   // polytope n-faces (sorted)
  int own_ptr[kMaxNf + 1], own_lst[kMaxNd]; // own DOFs per n-face (reference order)
  double qp[3][kMaxNq], qw[kMaxNq];
  double shp[kMaxNd][kMaxNq];               // scalar shape values
  double dshp[kMaxNd][3][kMaxNq];           // reference gradients of the scalar shapes (cell [-1, 1]^d)
  double gshp[8][kMaxNq];                   // geometric Q1 shape values
  double gdshp[8][3][kMaxNq];               // reference gradients of the geometric Q1 shapes
  double detj[kMaxNq];
  double grad[kMaxNd][3][kMaxNq];           // physical gradients of scalar shapes (first cell map)
};
void build_reffe(RefFE &fe, int dim, int order, int ncomp, const double h[3], int nq1 = 0);
void gauss_legendre(int n, double *x, double *w);

// --------------------------------------------------------------- context ---
}  // namespace femhip

// The context is reference counted: every space, operator and solver holds a
// reference, so femhip_ctx_destroy may be called before its dependents are
// destroyed (the stream is released with the last reference).
struct femhip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  std::atomic<int> refs{1};
};

namespace femhip {
inline femhip_ctx *ctx_retain(femhip_ctx *c) {
  c->refs.fetch_add(1);
  return c;
}
inline void ctx_release(femhip_ctx *c) {
  if (c && c->refs.fetch_sub(1) == 1) {
    (void)hipSetDevice(c->device);
    if (c->stream) {
      (void)hipStreamSynchronize(c->stream);
      (void)hipStreamDestroy(c->stream);
    }
    delete c;
  }
}
}  // namespace femhip

struct femhip_space {
  femhip_ctx *ctx = nullptr;
  std::atomic<int> refs{1};           // the handle + every operator built on it
  femhip_mesh_desc_t mesh{};
  femhip::RefFE fe;
  int64_t ncells = 0;
  int32_t n_free = 0, n_fixed = 0;
  std::vector<int32_t> dofs;          // host copy (cells x nd), signed 1-based; made on request
  femhip::DevBuf<int32_t> d_dofs;     // device: free -> row (0-based), fixed -> -(fixed index)-1
  femhip::DevBuf<int64_t> d_rep;      // device: free row -> cell * nd + local DOF of one occurrence
  // vertex coordinates (num_vertices x dim, x fastest) when the caller moved
  // the vertices: every cell then has its own trilinear map
  femhip::DevBuf<double> vcoords;
  bool mapped = false;
};

namespace femhip {
// Solver copy of the matrix in 64-row slices (one wavefront per slice).  A
// slice whose rows share few distinct column offsets (col - row) is stored by
// diagonals: K offsets per slice and K x 64 values, no column indices, and the
// x gathers of one diagonal are contiguous.  Other slices are stored as ELL
// with K = longest row and explicit column indices.  Values are laid out
// [slice][k][lane]; padding entries hold 0 and a valid column.  Where rows of
// very different lengths mix (Q2: 27 to 125 entries), the rows of a small
// window of consecutive rows may be sorted by length before they are cut into
// ELL slices, so a slice's rows have similar lengths (little padding) while
// its gathers stay local; such slices carry the row of each lane in `srow`.
constexpr int kSlice = 64;
constexpr int kSortWindow = 512;
// Position of stored diagonal k of row r in the symmetric-diagonal copy (K
// stored diagonals, rows padded to `stride`): diagonal-major, k * stride + r.
// (Slice-major [r / 64][k][r % 64], one contiguous block per slice, was
// measured slower: 59.3 against 54.3 us for the 128^3 FP64 SpMV, 273 against
// 247 us at 192^3.)
__host__ __device__ __forceinline__ int64_t sym_at(int64_t k, int64_t r, int64_t stride, int K) {
  (void)K;
  return k * stride + r;
}
//
// Symmetric-diagonal mode (sym = true): when every row's column offsets come
// from one global set D = -D (structured Q1 numbering), only the diagonals with
// offset >= 0 are kept, diagonal-major with row stride `stride`; the entry
// (r, r - o) of a lower diagonal is read as the stored (r - o, r), a shifted
// but still contiguous load.  The matrix stream is then about half of the full
// pattern's, as with FEMPAR's symmetric storage.
struct Sliced {
  int32_t ns = 0;                 // number of slices
  int64_t entries = 0;            // padded entries
  int64_t dia_slices = 0;
  int64_t perm_slices = 0;        // ELL slices over sorted windows
  bool sym = false;               // symmetric-diagonal mode
  int32_t kplus = 0;              // stored diagonals (offsets >= 0, first is 0)
  int64_t stride = 0;             // row stride of a stored diagonal (n rounded to 64)
  std::vector<int32_t> h_soffs;
  DevBuf<int32_t> soffs;          // kplus offsets, ascending
  DevBuf<int64_t> base;           // ns+1 entry offsets
  DevBuf<int32_t> meta;           // ns: K | (is_ell << 30) | (clamp << 29) | (permuted << 28)
  DevBuf<int32_t> srow;           // ns x 64: row of each lane of a permuted slice (n: none)
  DevBuf<int32_t> aux;            // ns: offset into offs (diagonal) or ecol (ELL)
  DevBuf<int32_t> offs;           // diagonal offsets
  DevBuf<int32_t> ecol;           // ELL column indices
  DevBuf<int32_t> map;            // entries: CSR position, -1 for padding (not used by the symmetric mode)
  DevBuf<double> v64;
  DevBuf<float> v32;
  int64_t filled_version = -1, filled_version32 = -1;
  bool scaled32 = false;          // v32 holds S A S (S = diag^(-1/2)) rather than A
  bool wants_int32 = false;       // v32 was asked for integrated in FP32 (FP32_STRICT)
  bool built = false;
  // Solver rows: the operator's rows, or (permuted) the class-blocked order of
  // a structured scalar space of order >= 2, padded with identity rows; perm
  // maps an operator row to its solver row.
  int32_t rows = 0;
  bool permuted = false;
  DevBuf<int32_t> perm;
};
}  // namespace femhip

struct femhip_operator {
  femhip_ctx *ctx = nullptr;
  std::atomic<int> refs{1};           // the handle + every solver built on it
  int64_t version = 0;                // bumped by every compute()
  femhip::Sliced sl;                  // solver format (built on first solve)
  femhip_space *space = nullptr;
  int symmetric_storage = 0;
  int32_t n = 0;
  int64_t nnz_full = 0, nnz_upper = 0;
  // full CSR (solver format), 0-based, built on the device; host copies are
  // made on request (pattern API, host-built solver formats)
  std::vector<int32_t> h_rowptr, h_col;
  femhip::DevBuf<int32_t> rowptr, col;
  femhip::DevBuf<double> val, rhs;
  femhip::DevBuf<float> val32;              // FP32 copy for the refined solver (lazy)
  // physics
  femhip_physics_t phys{};
  bool has_physics = false;
  femhip::DevBuf<double> k_cell;
  femhip::DevBuf<double> tab;               // element tables (per-q matrices etc.)
  int tab_kind = 0;
  bool computed = false;
  femhip::DevBuf<double> fixed;             // fixed values on device
  // owner-computes assembly of scalar Q1 (one row per free vertex): vertex ->
  // signed DOF id, free row -> vertex, per row the CSR slot of each of the 3^d
  // neighbour offsets (255: none), and per direction the sin/cos factors of
  // the coefficients at the 1-D quadrature points of every cell layer
  int gather = 0;                           // 0 not set up, 1 in use, -1 not applicable
  femhip::DevBuf<int32_t> vtx_dof, row_vtx;
  femhip::DevBuf<uint8_t> slot;
  femhip::DevBuf<double> trig;              // 4 values per (direction, layer, 1-D point)
  // 3-D Q2: every cell's element matrix (27 x 27, row-major) and vector,
  // formed on the matrix cores, then gathered by the rows' owners
  femhip::DevBuf<double> elmat, elvec;
  // per-cell integration (cells with their own map, coefficients at the
  // quadrature points): reference-FE table and the caller's point values
  femhip::DevBuf<double> mtab, k_qp, f_qp;
  int trig_off[3] = {0, 0, 0};
  // the owner-computes path also writes the solver's symmetric-diagonal copy
  // when every neighbour code has one column offset over all rows: sym_kidx
  // [code] = stored diagonal of that offset (-1: lower part).  sym_written:
  // copies the last compute wrote (1: FP64; the FP32 copy is derived from it).
  int sym_gather = 0;                       // 0 not checked, 1 in use, -1 not applicable
  int8_t sym_kidx[32];
  int sym_written = 0;
  // structured stencil operator (no CSR, no space): rows are component-major
  // blocks of `blk` rows (the `nodes` free nodes, padded to a multiple of 64),
  // values assembled in place into the sliced layout
  bool structured = false;
  int32_t blk = 0, nodes = 0;
};

namespace femhip {

// n-face tables of a space's reference FE and its mesh, passed to kernels by
// value: per n-face its type (tangent axes), anchor, own DOFs; per local DOF
// its n-face and position in that face's own list; (type, anchor) -> n-face
struct FaceTab {
  int32_t dim, nd, nnf, fc, bc, np1, nnode, order;
  int32_t n[3];
  int32_t type[kMaxNf], anchor[kMaxNf], cnt[kMaxNf], ptr[kMaxNf + 1];
  int8_t lst[kMaxNd];
  int8_t face_of[kMaxNd];
  int8_t pos_of[kMaxNd];
  int8_t face_by[8][8];
  double lo[3], hi[3];
};
FaceTab face_tab(const femhip_space &s);

// (cell counts are below 2^31: 32-bit division)
__device__ __forceinline__ void cell_ijk(const FaceTab &t, int64_t c, int (&ijk)[3]) {
  int v = (int)c;
  ijk[0] = v % t.n[0];
  v /= t.n[0];
  ijk[1] = v % t.n[1];
  ijk[2] = v / t.n[1];
}

__device__ __forceinline__ int64_t cell_index(const FaceTab &t, const int (&ijk)[3]) {
  return ijk[0] + (int64_t)t.n[0] * (ijk[1] + (int64_t)t.n[1] * ijk[2]);
}

// device set-up (setup.hip): first-touch numbering of a structured space
// (d_dofs, d_rep, n_free, n_fixed); representatives of a caller's numbering,
// checked to conform to the mesh; the CSR pattern on the device (rowptr, col,
// n, nnz_full, nnz_upper); host copies on request; the symmetric-diagonal
// test (every row's column offsets in the longest row's set D, structurally
// symmetric); the linear Dirichlet function at the fixed DOFs
void number_structured(femhip_space &s);
// exclusive scan of n counts on the stream (out may alias in); out[n] = total
// when with_total; returns the total (checked below 2^31)
int64_t exclusive_scan(hipStream_t st, const int32_t *in, int32_t *out, int64_t n, bool with_total);
void representatives_from_dofs(femhip_space &s);
void build_pattern(femhip_operator &op);
void ensure_host_pattern(const femhip_operator &op);
void ensure_host_dofs(const femhip_space &s);
bool symmetric_offsets(const femhip_operator &op, std::vector<int32_t> &D);
void interpolate_dirichlet_linear(const femhip_space &s, double *out_host);
// sliced solver format (sliced.cpp)
void build_sliced(femhip_operator &op);

// assembly (assemble.hip)
void assemble(femhip_operator &op, const double *fixed_host, float *ms);
// once the symmetric-diagonal solver copy exists: decide whether the
// owner-computes assembly can write it directly (sets op.sym_gather)
void setup_sym_gather(femhip_operator &op);
// FP32 symmetric-diagonal copy of A, or of S A S with sc = S, integrated in
// FP32 on the matrix cores (scalar Q1); false where that does not apply
bool assemble_sym32(femhip_operator &op, const double *sc, float *v32);
void upload_physics(femhip_operator &op);
void spmv_full(femhip_operator &op, const double *d_x, double *d_y);
// address of the no-allocation staging array on `device` (noalloc.hip); called
// by femhip_ctx_create with that device current
void noalloc_setup(int device);

// per-cell integration (mapped.hip): whether an operator's cells need it (the
// space's vertices were moved, or k or f are given at the quadrature points);
// the space's vertex coordinates; the physical quadrature points of every
// cell; linear Dirichlet values at the mapped nodes; the caller's point values
// copied to the device; every cell's element matrix and vector (op.elmat,
// op.elvec); error norms with each cell's own map
bool general_integration(const femhip_operator &op);
void set_vertex_coords(femhip_space &s, const double *coords);
void quadrature_points(const femhip_space &s, double *xq_host);
void interpolate_dirichlet_mapped(const femhip_space &s, double *out_host);
void upload_qpoint_values(femhip_operator &op);
void element_matrices_mapped(femhip_operator &op);
void error_norms_mapped(femhip_operator &op, const double *d_x, int exact, double out[4]);

// element matrix (row-major) and body-force vector of the elasticity plugin
void elasticity_element(const RefFE &fe, const femhip_physics_t &ph, double *Ke, double *Fe);

// Structured elasticity operator on the unit cube with n cells per direction,
// clamped on x = 0 (stencil.hip): numbering, solver format, values, RHS built
// on the device of ctx; row (c, node) = c * blk + node, node lexicographic
// over the free vertices (x index 1..n fastest, then y, then z).
femhip_operator *create_structured_elasticity(femhip_ctx *ctx, int32_t n, const femhip_physics_t &ph);
// whether create_structured_elasticity accepts n (its sliced format indexes
// 243 x padded vertices entries with 32-bit integers)
bool structured_elasticity_fits(int32_t n);

// error norms of the FE function with free values d_x (norms.hip):
// out = {l2_norm, h1_seminorm, h1_norm, linfty_norm}
void error_norms(femhip_operator &op, const double *d_x, int exact, double out[4]);

// non-blocking CG driver (cg.hip) used by the request server: solver_start
// prepares the solver format and enqueues the initial residual for b = the
// operator's RHS and x0 = 0; solver_push enqueues one chunk of solver_chunk()
// iterations whose last kernel writes the result block into pinned slot 0 or 1;
// solver_poll returns 0 while that chunk is running, else 2 once the device
// state says done and 1 otherwise; solver_finish waits for the stream.
// solver_stream_launches makes the solver enqueue its chunks as plain stream
// launches instead of hipGraphs (set before solver_start).
int solver_chunk();
void solver_stream_launches(femhip_solver &s);
// Pinned, device-mapped result slots and completion events a solver may use
// instead of allocating its own (the request server keeps one set per slot
// and reuses it for every request of that slot).
struct SolverHostSlots {
  void *h_res[2] = {nullptr, nullptr};   // CgResult each
  int *h_flag = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
};
void solver_use_host_slots(femhip_solver &s, const SolverHostSlots &hs);
size_t solver_result_bytes();
void solver_start(femhip_solver &s);
void solver_push(femhip_solver &s, int slot);
int solver_poll(femhip_solver &s, int slot);
void solver_finish(femhip_solver &s, femhip_cg_result_t *out);
void solver_copy_solution(const femhip_solver &s, double *x);
size_t solver_device_bytes(const femhip_solver &s);

// device memory held by an operator (CSR, solver format, tables) and its space
inline size_t operator_device_bytes(const femhip_operator &op) {
  size_t b = op.rowptr.bytes() + op.col.bytes() + op.val.bytes() + op.rhs.bytes() + op.val32.bytes() +
             op.k_cell.bytes() + op.tab.bytes() + op.fixed.bytes() + op.vtx_dof.bytes() +
             op.row_vtx.bytes() + op.slot.bytes() + op.trig.bytes() + op.elmat.bytes() + op.elvec.bytes() +
             op.mtab.bytes() + op.k_qp.bytes() + op.f_qp.bytes();
  const Sliced &s = op.sl;
  b += s.perm.bytes() + s.soffs.bytes() + s.base.bytes() + s.meta.bytes() + s.srow.bytes() + s.aux.bytes() + s.offs.bytes() +
       s.ecol.bytes() +
       s.map.bytes() + s.v64.bytes() + s.v32.bytes();
  if (op.space) b += op.space->d_dofs.bytes() + op.space->d_rep.bytes() + op.space->vcoords.bytes();
  return b;
}

}  // namespace femhip
