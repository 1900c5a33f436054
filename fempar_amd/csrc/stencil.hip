// stencil.hip -- structured elasticity operator for the request server.
//
// The request server's problems are vector-Q1 linear elasticity on the unit
// cube with n cells per direction, all components clamped on x = 0 and
// constant (E, nu, body force).  Every cell has the same element matrix (the
// irreducible plugin, sbm_irreducible_discrete_integration.i90:64-130), so the
// operator is a 27-point block stencil whose boundary rows keep only the cells
// that exist.  It is built on the device without DOF lists, CSR pattern or
// host passes over the matrix:
//   * rows are component-major: row (c, node) = c * blk + node, with the free
//     vertices (x index 1..n) numbered x-fastest and each component block
//     padded to a multiple of 64 rows (padding rows are identity rows, RHS 0);
//   * every 64-row slice lies in one component block, so all its rows share
//     the 81 column offsets (c' - c) blk + dx + mx dy + mx my dz: the sliced
//     solver format holds diagonal slices only, with no column indices;
//   * one thread per row sums, over the cells around its vertex in ascending
//     cell order, the element entries coupling it to each neighbour DOF and
//     writes the row's FP64 and FP32 values and RHS entry once.
#include <hip/hip_runtime.h>

#include <vector>

#include "femhip_internal.h"

namespace femhip {
namespace {

constexpr int kK = 81;   // 3 components x 27 neighbour vertices

struct StencilDev {
  int n;      // cells per direction
  int blk;    // rows per component block (multiple of 64)
  int nodes;  // free vertices, n (n+1)^2
};

__global__ __launch_bounds__(256) void k_stencil_assemble(StencilDev g, const double *__restrict__ Ke,
                                                          const double *__restrict__ Fe, double *__restrict__ v64,
                                                          float *__restrict__ v32, double *__restrict__ rhs) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= 3 * (int64_t)g.blk) return;
  const int c = (int)(r / g.blk);   // uniform over a wavefront: blk is a multiple of 64
  const int node = (int)(r % g.blk);
  const int64_t s = r / 64, lane = r % 64;
  double *V = v64 + s * (int64_t)kK * 64 + lane;
  float *V32 = v32 + s * (int64_t)kK * 64 + lane;
  if (node >= g.nodes) {   // padding row: identity, zero RHS
    for (int k = 0; k < kK; ++k) {
      const double v = k == c * 27 + 13 ? 1.0 : 0.0;
      V[k * 64] = v;
      V32[k * 64] = (float)v;
    }
    rhs[r] = 0.0;
    return;
  }
  const int n = g.n, mx = n, my = n + 1;
  const int a = node % mx + 1, b = (node / mx) % my, z = node / (mx * my);   // vertex, a in 1..n
  double acc[kK];
#pragma unroll
  for (int k = 0; k < kK; ++k) acc[k] = 0.0;
  double bacc = 0.0;
#pragma unroll
  for (int e = 0; e < 8; ++e) {   // cells around the vertex, ascending cell index
    const int ex = e & 1, ey = (e >> 1) & 1, ez = e >> 2;
    const int ci = a - 1 + ex, cj = b - 1 + ey, ck = z - 1 + ez;
    if (ci < 0 || ci >= n || cj < 0 || cj >= n || ck < 0 || ck >= n) continue;
    const int il = (1 - ex) + 2 * (1 - ey) + 4 * (1 - ez);   // the vertex in this cell
    const double *Krow = Ke + (c * 8 + il) * 24;              // local DOF index comp * 8 + vertex
    bacc += Fe[c * 8 + il];
#pragma unroll
    for (int jl = 0; jl < 8; ++jl) {
      const int dx = ex + (jl & 1) - 1, dy = ey + ((jl >> 1) & 1) - 1, dz = ez + (jl >> 2) - 1;
      if (a + dx < 1) continue;   // neighbour on the clamped face: fixed with g = 0
      const int code = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
#pragma unroll
      for (int c2 = 0; c2 < 3; ++c2) acc[c2 * 27 + code] += Krow[c2 * 8 + jl];
    }
  }
#pragma unroll
  for (int k = 0; k < kK; ++k) {
    V[k * 64] = acc[k];
    V32[k * 64] = (float)acc[k];
  }
  rhs[r] = bacc;
}

}  // namespace

bool structured_elasticity_fits(int32_t n) {
  if (n < 1 || n > 1024) return false;
  const int64_t nodes = (int64_t)n * (n + 1) * (n + 1);
  const int64_t blk = (nodes + 63) / 64 * 64;
  return 3 * blk * kK < (int64_t)INT32_MAX;
}

femhip_operator *create_structured_elasticity(femhip_ctx *ctx, int32_t n, const femhip_physics_t &ph) {
  FH_CHECK(structured_elasticity_fits(n), FEMHIP_ERR_INVALID,
           "structured elasticity needs 1 <= n with 243 x padded vertices < 2^31 (n <= 206)");
  FH_CHECK(ph.kind == FEMHIP_PHYS_ELASTICITY && ph.E > 0 && ph.nu > -1.0 && ph.nu < 0.5, FEMHIP_ERR_INVALID,
           "invalid elasticity parameters");
  const int64_t nodes = (int64_t)n * (n + 1) * (n + 1);
  const int64_t blk = (nodes + 63) / 64 * 64;
  const int64_t rows = 3 * blk;
  const int64_t entries = rows * kK;
  auto *op = new femhip_operator();
  op->ctx = ctx_retain(ctx);
  try {
    hipStream_t st = ctx->stream;
    op->structured = true;
    op->blk = (int32_t)blk;
    op->nodes = (int32_t)nodes;
    op->n = (int32_t)rows;
    op->phys = ph;
    op->has_physics = true;
    // element matrix of the first cell (uniform spacing 1/n)
    RefFE fe;
    const double h[3] = {1.0 / n, 1.0 / n, 1.0 / n};
    build_reffe(fe, 3, 1, 3, h);
    std::vector<double> t(24 * 24 + 24);
    elasticity_element(fe, ph, t.data(), t.data() + 24 * 24);
    op->tab.alloc(t.size());
    FH_HIP(hipMemcpyAsync(op->tab.p, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, st));
    // sliced format: one set of 81 offsets per component block
    const int32_t mx = n, my = n + 1;
    std::vector<int32_t> offs(3 * kK);
    for (int c = 0; c < 3; ++c)
      for (int c2 = 0; c2 < 3; ++c2)
        for (int code = 0; code < 27; ++code) {
          const int dx = code % 3 - 1, dy = (code / 3) % 3 - 1, dz = code / 9 - 1;
          offs[(size_t)(c * kK + c2 * 27 + code)] = (c2 - c) * (int32_t)blk + dx + mx * (dy + my * dz);
        }
    Sliced &S = op->sl;
    S.ns = (int32_t)(rows / 64);
    S.entries = entries;
    S.dia_slices = S.ns;
    S.sym = false;
    std::vector<int64_t> base((size_t)S.ns + 1);
    std::vector<int32_t> meta((size_t)S.ns), aux((size_t)S.ns);
    for (int32_t s = 0; s < S.ns; ++s) {
      const int c = (int)((int64_t)s * 64 / blk);
      base[(size_t)s] = (int64_t)s * kK * 64;
      aux[(size_t)s] = c * kK;
      meta[(size_t)s] = kK;
      // bit 29: some lane's column leaves [0, rows) (the values there are 0)
      const int64_t lo = (int64_t)s * 64 + offs[(size_t)(c * kK)];
      const int64_t hi = (int64_t)s * 64 + 63 + offs[(size_t)(c * kK + kK - 1)];
      if (lo < 0 || hi >= rows) meta[(size_t)s] |= 1 << 29;
    }
    base[(size_t)S.ns] = entries;
    S.base.alloc(base.size());
    S.meta.alloc(meta.size());
    S.aux.alloc(aux.size());
    S.offs.alloc(offs.size());
    S.ecol.alloc(1);
    S.map.alloc(1);
    S.v64.alloc((size_t)entries);
    S.v32.alloc((size_t)entries);
    op->rhs.alloc((size_t)rows);
    FH_HIP(hipMemcpyAsync(S.base.p, base.data(), sizeof(int64_t) * base.size(), hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(S.meta.p, meta.data(), sizeof(int32_t) * meta.size(), hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(S.aux.p, aux.data(), sizeof(int32_t) * aux.size(), hipMemcpyHostToDevice, st));
    FH_HIP(hipMemcpyAsync(S.offs.p, offs.data(), sizeof(int32_t) * offs.size(), hipMemcpyHostToDevice, st));
    const StencilDev g{n, (int)blk, (int)nodes};
    hipLaunchKernelGGL(k_stencil_assemble, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st, g, op->tab.p,
                       op->tab.p + 24 * 24, S.v64.p, S.v32.p, op->rhs.p);
    FH_HIP(hipGetLastError());
    // the host vectors above must outlive the copies
    FH_HIP(hipStreamSynchronize(st));
    S.built = true;
    op->computed = true;
    op->version = 1;
    S.filled_version = S.filled_version32 = 1;   // values were assembled in place
  } catch (...) {
    femhip_operator_destroy(op);
    throw;
  }
  return op;
}

}  // namespace femhip
