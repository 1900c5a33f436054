// api.cpp -- C-ABI entry points for context, FE space and operator.
#include <cmath>
#include <cstring>

#include "femhip_internal.h"

namespace femhip {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
DevArena *&current_arena() {
  static thread_local DevArena *a = nullptr;
  return a;
}
}  // namespace femhip

using namespace femhip;

#define FH_TRY try {
#define FH_CATCH                 \
  }                              \
  catch (const Error &e) {       \
    set_error(e.msg);            \
    return e.code;               \
  }                              \
  catch (const std::bad_alloc &) { \
    set_error("host allocation failed"); \
    return FEMHIP_ERR_OOM;       \
  }

namespace {

void check_mesh(const femhip_mesh_desc_t *m) {
  FH_CHECK(m, FEMHIP_ERR_INVALID, "null mesh descriptor");
  FH_CHECK(m->dim == 2 || m->dim == 3, FEMHIP_ERR_INVALID, "dim must be 2 or 3");
  for (int k = 0; k < m->dim; ++k) {
    FH_CHECK(m->n[k] >= 1, FEMHIP_ERR_INVALID, "num_cells_x_dim must be >= 1");
    FH_CHECK(m->limits[2 * k + 1] > m->limits[2 * k], FEMHIP_ERR_INVALID, "domain limits must be increasing");
  }
  FH_CHECK(m->order == 1 || m->order == 2, FEMHIP_ERR_INVALID, "Lagrangian order must be 1 or 2");
  FH_CHECK(m->num_components == 1 || m->num_components == m->dim, FEMHIP_ERR_INVALID,
           "num_components must be 1 or dim");
  FH_CHECK(!(m->num_components > 1 && m->order > 1), FEMHIP_ERR_INVALID, "vector fields support order 1");
  FH_CHECK(m->dirichlet >= 0 && m->dirichlet <= 2, FEMHIP_ERR_INVALID, "unknown Dirichlet kind");
  int64_t cells = 1;
  for (int k = 0; k < m->dim; ++k) cells *= m->n[k];
  FH_CHECK(cells < (int64_t)1 << 31, FEMHIP_ERR_INVALID, "too many cells for 32-bit ids");
}

void init_space(femhip_space &s, femhip_ctx *ctx, const femhip_mesh_desc_t *mesh) {
  s.ctx = ctx_retain(ctx);
  s.mesh = *mesh;
  if (s.mesh.dim == 2) {
    s.mesh.n[2] = 1;
    s.mesh.limits[4] = 0.0;
    s.mesh.limits[5] = 0.0;
  }
  // spacing of the first cell, as the reference's first cell map sees it
  double h[3] = {1.0, 1.0, 1.0};
  for (int k = 0; k < s.mesh.dim; ++k) {
    const double a = s.mesh.limits[2 * k], b = s.mesh.limits[2 * k + 1];
    h[k] = ((b - a) * (1.0 / (double)s.mesh.n[k]) + a) - a;
  }
  build_reffe(s.fe, s.mesh.dim, s.mesh.order, s.mesh.num_components, h);
}

void upload_dofs(femhip_space &s) {
  const size_t cnt = (size_t)(s.ncells * s.fe.nd);
  std::vector<int32_t> dev(cnt);
  for (size_t i = 0; i < cnt; ++i) {
    int32_t g = s.dofs[i];
    dev[i] = g > 0 ? g - 1 : g;   // free -> 0-based row, fixed keeps its negative id
  }
  s.d_dofs.alloc(cnt);
  FH_HIP(hipMemcpy(s.d_dofs.p, dev.data(), cnt * sizeof(int32_t), hipMemcpyHostToDevice));
}

}  // namespace

extern "C" {

const char *femhip_last_error(void) { return g_last_error.c_str(); }

int femhip_ctx_create(int32_t device, femhip_ctx_t **ctx) {
  FH_TRY
  FH_CHECK(ctx, FEMHIP_ERR_INVALID, "null ctx pointer");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    throw Error{FEMHIP_ERR_NO_DEVICE, "no HIP device available: libfemhip has no CPU fallback"};
  FH_CHECK(device >= 0 && device < ndev, FEMHIP_ERR_INVALID, "device index out of range");
  FH_HIP(hipSetDevice(device));
  auto *c = new femhip_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  FH_HIP(hipGetDeviceProperties(&prop, device));
  c->num_cus = prop.multiProcessorCount;
  try {
    noalloc_setup(device);
    FH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  } catch (...) {
    delete c;
    throw;
  }
  *ctx = c;
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_ctx_destroy(femhip_ctx_t *ctx) {
  ctx_release(ctx);   // freed with the last dependent handle
  return FEMHIP_OK;
}

int femhip_space_create_structured(femhip_ctx_t *ctx, const femhip_mesh_desc_t *mesh, femhip_space_t **space) {
  FH_TRY
  FH_CHECK(ctx && space, FEMHIP_ERR_INVALID, "null argument");
  check_mesh(mesh);
  FH_HIP(hipSetDevice(ctx->device));
  auto *s = new femhip_space();
  try {
    init_space(*s, ctx, mesh);
    number_structured(*s);
  } catch (...) {
    femhip_space_destroy(s);
    throw;
  }
  *space = s;
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_space_create_from_dofs(femhip_ctx_t *ctx, const femhip_mesh_desc_t *mesh, const int32_t *lst_dofs_gids,
                                  int32_t n_free, int32_t n_fixed, femhip_space_t **space) {
  FH_TRY
  FH_CHECK(ctx && space && lst_dofs_gids, FEMHIP_ERR_INVALID, "null argument");
  check_mesh(mesh);
  FH_CHECK(n_free >= 0 && n_fixed >= 0, FEMHIP_ERR_INVALID, "negative DOF counts");
  FH_HIP(hipSetDevice(ctx->device));
  auto *s = new femhip_space();
  try {
    init_space(*s, ctx, mesh);
    s->ncells = 1;
    for (int k = 0; k < s->mesh.dim; ++k) s->ncells *= s->mesh.n[k];
    const size_t cnt = (size_t)(s->ncells * s->fe.nd);
    s->dofs.assign(lst_dofs_gids, lst_dofs_gids + cnt);
    for (size_t i = 0; i < cnt; ++i) {
      int32_t g = s->dofs[i];
      FH_CHECK(g != 0 && g <= n_free && -g <= n_fixed, FEMHIP_ERR_INVALID,
               "lst_dofs_gids entry out of range [-n_fixed, n_free] or zero");
    }
    s->n_free = n_free;
    s->n_fixed = n_fixed;
    upload_dofs(*s);
    representatives_from_dofs(*s);
  } catch (...) {
    femhip_space_destroy(s);
    throw;
  }
  *space = s;
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_space_info(const femhip_space_t *s, int64_t *num_cells, int32_t *nd, int32_t *nq, int32_t *n_free,
                      int32_t *n_fixed) {
  FH_TRY
  FH_CHECK(s, FEMHIP_ERR_INVALID, "null space");
  if (num_cells) *num_cells = s->ncells;
  if (nd) *nd = s->fe.nd;
  if (nq) *nq = s->fe.nq;
  if (n_free) *n_free = s->n_free;
  if (n_fixed) *n_fixed = s->n_fixed;
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_space_get_dofs(const femhip_space_t *s, int32_t *out) {
  FH_TRY
  FH_CHECK(s && out, FEMHIP_ERR_INVALID, "null argument");
  FH_HIP(hipSetDevice(s->ctx->device));
  ensure_host_dofs(*s);
  std::memcpy(out, s->dofs.data(), s->dofs.size() * sizeof(int32_t));
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_space_interpolate_dirichlet(const femhip_space_t *s, const femhip_physics_t *phys, double *out) {
  FH_TRY
  FH_CHECK(s && phys && (out || s->n_fixed == 0), FEMHIP_ERR_INVALID, "null argument");
  for (int32_t i = 0; i < s->n_fixed; ++i) out[i] = 0.0;
  if (phys->g_kind == FEMHIP_G_ZERO) return FEMHIP_OK;
  FH_CHECK(phys->g_kind == FEMHIP_G_LINEAR, FEMHIP_ERR_INVALID, "unknown Dirichlet function");
  FH_HIP(hipSetDevice(s->ctx->device));
  // g(x) = sum of the coordinates at each fixed node
  if (s->mapped) interpolate_dirichlet_mapped(*s, out);
  else interpolate_dirichlet_linear(*s, out);
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_space_set_vertex_coords(femhip_space_t *s, const double *coords) {
  FH_TRY
  FH_CHECK(s && coords, FEMHIP_ERR_INVALID, "null argument");
  FH_HIP(hipSetDevice(s->ctx->device));
  set_vertex_coords(*s, coords);
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_space_quadrature_points(const femhip_space_t *s, double *xq) {
  FH_TRY
  FH_CHECK(s && xq, FEMHIP_ERR_INVALID, "null argument");
  FH_HIP(hipSetDevice(s->ctx->device));
  quadrature_points(*s, xq);
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_space_destroy(femhip_space_t *s) {
  if (s && s->refs.fetch_sub(1) == 1) {
    femhip_ctx *c = s->ctx;
    if (c) (void)hipSetDevice(c->device);
    delete s;
    ctx_release(c);
  }
  return FEMHIP_OK;
}

int femhip_operator_create(femhip_ctx_t *ctx, femhip_space_t *space, int32_t symmetric_storage,
                           femhip_operator_t **op) {
  FH_TRY
  FH_CHECK(ctx && space && op, FEMHIP_ERR_INVALID, "null argument");
  FH_HIP(hipSetDevice(ctx->device));
  auto *o = new femhip_operator();
  o->ctx = ctx_retain(ctx);
  space->refs.fetch_add(1);
  o->space = space;
  try {
    o->symmetric_storage = symmetric_storage ? 1 : 0;
    build_pattern(*o);   // rowptr, col on the device
    o->val.alloc((size_t)o->nnz_full);
    o->rhs.alloc((size_t)std::max(1, o->n));
  } catch (...) {
    delete o;
    femhip_space_destroy(space);
    ctx_release(ctx);
    throw;
  }
  *op = o;
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_operator_info(const femhip_operator_t *op, int32_t *n, int64_t *nnz, int64_t *nnz_full) {
  FH_TRY
  FH_CHECK(op, FEMHIP_ERR_INVALID, "null operator");
  if (n) *n = op->n;
  if (nnz) *nnz = op->symmetric_storage ? op->nnz_upper : op->nnz_full;
  if (nnz_full) *nnz_full = op->nnz_full;
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_operator_get_pattern(const femhip_operator_t *op, int32_t *irp, int32_t *ja) {
  FH_TRY
  FH_CHECK(op && irp && ja, FEMHIP_ERR_INVALID, "null argument");
  FH_HIP(hipSetDevice(op->ctx->device));
  ensure_host_pattern(*op);
  if (!op->symmetric_storage) {
    for (int32_t r = 0; r <= op->n; ++r) irp[r] = op->h_rowptr[(size_t)r] + 1;
    for (int64_t k = 0; k < op->nnz_full; ++k) ja[k] = op->h_col[(size_t)k] + 1;
  } else {
    int64_t u = 0;
    irp[0] = 1;
    for (int32_t r = 0; r < op->n; ++r) {
      for (int64_t k = op->h_rowptr[(size_t)r]; k < op->h_rowptr[(size_t)r + 1]; ++k)
        if (op->h_col[(size_t)k] >= r) ja[u++] = op->h_col[(size_t)k] + 1;
      irp[r + 1] = (int32_t)(u + 1);
    }
  }
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_operator_set_physics(femhip_operator_t *op, const femhip_physics_t *phys) {
  FH_TRY
  FH_CHECK(op && phys, FEMHIP_ERR_INVALID, "null argument");
  FH_CHECK(phys->kind == FEMHIP_PHYS_POISSON || phys->kind == FEMHIP_PHYS_ELASTICITY, FEMHIP_ERR_INVALID,
           "unknown physics");
  FH_CHECK(phys->k_kind >= 0 && phys->k_kind <= 3 && phys->f_kind >= 0 && phys->f_kind <= 3, FEMHIP_ERR_INVALID,
           "unknown coefficient/source kind");
  if (phys->kind == FEMHIP_PHYS_ELASTICITY)
    FH_CHECK(phys->E > 0 && phys->nu > -1.0 && phys->nu < 0.5, FEMHIP_ERR_INVALID, "invalid E or nu");
  FH_HIP(hipSetDevice(op->ctx->device));
  op->phys = *phys;
  op->has_physics = true;
  upload_physics(*op);
  op->computed = false;
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_operator_compute(femhip_operator_t *op, const double *fixed_values, double *ms) {
  FH_TRY
  FH_CHECK(op, FEMHIP_ERR_INVALID, "null operator");
  FH_CHECK(op->has_physics, FEMHIP_ERR_STATE, "set_physics (the discrete_integration plugin) not called");
  FH_HIP(hipSetDevice(op->ctx->device));
  float t = 0.f;
  assemble(*op, fixed_values, &t);
  op->computed = true;
  ++op->version;
  if (op->sym_written & 1) op->sl.filled_version = op->version;   // the assembly wrote the solver copy
  if (ms) *ms = t;
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_operator_get_values(const femhip_operator_t *op, double *val, double *rhs) {
  FH_TRY
  FH_CHECK(op, FEMHIP_ERR_INVALID, "null operator");
  FH_CHECK(op->computed, FEMHIP_ERR_STATE, "operator not computed");
  FH_HIP(hipSetDevice(op->ctx->device));
  FH_HIP(hipStreamSynchronize(op->ctx->stream));
  if (val) {
    std::vector<double> full((size_t)op->nnz_full);
    FH_HIP(hipMemcpy(full.data(), op->val.p, sizeof(double) * full.size(), hipMemcpyDeviceToHost));
    if (!op->symmetric_storage) {
      std::memcpy(val, full.data(), sizeof(double) * full.size());
    } else {   // entries on or above the diagonal, row by row
      ensure_host_pattern(*op);
      int64_t u = 0;
      for (int32_t r = 0; r < op->n; ++r)
        for (int64_t k = op->h_rowptr[(size_t)r]; k < op->h_rowptr[(size_t)r + 1]; ++k)
          if (op->h_col[(size_t)k] >= r) val[u++] = full[(size_t)k];
    }
  }
  if (rhs) FH_HIP(hipMemcpy(rhs, op->rhs.p, sizeof(double) * (size_t)op->n, hipMemcpyDeviceToHost));
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_operator_apply(femhip_operator_t *op, const double *x, double *y) {
  FH_TRY
  FH_CHECK(op && x && y, FEMHIP_ERR_INVALID, "null argument");
  FH_CHECK(op->computed, FEMHIP_ERR_STATE, "operator not computed");
  FH_HIP(hipSetDevice(op->ctx->device));
  DevBuf<double> dx, dy;
  dx.alloc((size_t)op->n);
  dy.alloc((size_t)op->n);
  FH_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * (size_t)op->n, hipMemcpyHostToDevice, op->ctx->stream));
  spmv_full(*op, dx.p, dy.p);
  FH_HIP(hipMemcpyAsync(y, dy.p, sizeof(double) * (size_t)op->n, hipMemcpyDeviceToHost, op->ctx->stream));
  FH_HIP(hipStreamSynchronize(op->ctx->stream));
  return FEMHIP_OK;
  FH_CATCH
}

int femhip_operator_destroy(femhip_operator_t *op) {
  if (op && op->refs.fetch_sub(1) == 1) {
    femhip_ctx *c = op->ctx;
    femhip_space *sp = op->space;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete op;
    femhip_space_destroy(sp);
    ctx_release(c);
  }
  return FEMHIP_OK;
}

}  // extern "C"
