// space.cpp -- host-side setup of the FE space and CSR pattern.
//
// Reference FE tables (quadrature, Lagrange shape functions, affine map of the
// first cell), first-touch DOF numbering on a structured hex mesh and the CSR
// pattern that FEMPAR's COO sort_and_compress produces.  Setup only: the timed
// hot path (integration, assembly, CG) runs on the GPU.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include "femhip_internal.h"

namespace femhip {

static int popc(int v) { return __builtin_popcount((unsigned)v); }

// Gauss-Legendre rule on [-1,1], ascending, weights renormalised to sum 2 as
// legendre_compute_glr does (ReferenceFE/sbm_quadrature.i90:857-862).
void gauss_legendre(int n, double *x, double *w) {
  const double pi = 3.14159265358979323846;
  std::vector<double> der(n);
  auto legendre = [n](double z, double &pn, double &dpn) {
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    pn = (n == 0) ? 1.0 : p1;
    double pm = (n <= 1) ? 1.0 : p0;
    dpn = n * (z * pn - pm) / (z * z - 1.0);
  };
  for (int i = 0; i < n; ++i) {
    double z = -std::cos(pi * (i + 0.75) / (n + 0.5)), pn, dpn;
    for (int it = 0; it < 100; ++it) {
      legendre(z, pn, dpn);
      double dz = pn / dpn;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    if ((n & 1) && i == n / 2) z = 0.0;
    legendre(z, pn, dpn);
    x[i] = z;
    der[i] = std::fabs(dpn);
  }
  for (int i = 0; i < n / 2; ++i) {
    double a = 0.5 * (x[n - 1 - i] - x[i]);
    x[i] = -a;
    x[n - 1 - i] = a;
    double b = 0.5 * (der[i] + der[n - 1 - i]);
    der[i] = der[n - 1 - i] = b;
  }
  double sum = 0.0;
  for (int i = 0; i < n; ++i) {
    w[i] = 2.0 / (1.0 - x[i]) / (1.0 + x[i]) / der[i] / der[i];
    sum += w[i];
  }
  for (int i = 0; i < n; ++i) w[i] = 2.0 * w[i] / sum;
}

// 1-D Lagrange polynomial through equispaced nodes of [-1,1]
// (ReferenceFE/sbm_polynomial.i90:61-111): value and first derivative.
static void lagrange(int order, int i, double x, double &v, double &dv) {
  double t[8];
  for (int k = 0; k <= order; ++k) t[k] = order > 0 ? (2.0 / order) * k - 1.0 : 0.0;
  double c = 1.0;
  for (int k = 0; k <= order; ++k)
    if (k != i) c *= (t[i] - t[k]);
  c = 1.0 / c;
  double p0 = 1.0, p1 = 0.0;
  for (int k = 0; k <= order; ++k) {
    if (k == i) continue;
    p1 = p1 * (x - t[k]) + p0;
    p0 = p0 * (x - t[k]);
  }
  v = p0 * c;
  dv = p1 * c;
}

void build_reffe(RefFE &fe, int dim, int order, int ncomp, const double h[3]) {
  std::memset(&fe, 0, sizeof(fe));
  fe.dim = dim;
  fe.order = order;
  fe.ncomp = ncomp;
  // polytope n-faces sorted by (dimension, bit index) (ReferenceFE/sbm_polytope.i90:175-217)
  std::vector<std::pair<int, int>> nf;   // (type, anchor)
  for (int t = 0; t < (1 << dim); ++t)
    for (int a = 0; a < (1 << dim); ++a)
      if ((a & t) == 0) nf.push_back({t, a});
  std::stable_sort(nf.begin(), nf.end(), [dim](const std::pair<int, int> &A, const std::pair<int, int> &B) {
    int da = popc(A.first), db = popc(B.first);
    if (da != db) return da < db;
    return ((A.first << dim) | A.second) < ((B.first << dim) | B.second);
  });
  fe.nnf = (int)nf.size();
  for (int i = 0; i < fe.nnf; ++i) {
    fe.nf_type[i] = nf[i].first;
    fe.nf_anchor[i] = nf[i].second;
  }
  fe.np1 = order + 1;
  fe.nnode = 1;
  for (int k = 0; k < dim; ++k) fe.nnode *= fe.np1;
  fe.nd = fe.nnode * ncomp;
  // own DOFs of each n-face, component-major (sbm_lagrangian_reference_fe.i90:1250-1286)
  std::vector<int> node_face(fe.nnode);
  for (int a = 0; a < fe.nnode; ++a) {
    int r = a, typ = 0, anc = 0;
    for (int k = 0; k < dim; ++k) {
      int m = r % fe.np1;
      r /= fe.np1;
      if (m > 0 && m < order) typ |= 1 << k;
      else if (m == order) anc |= 1 << k;
    }
    for (int f = 0; f < fe.nnf; ++f)
      if (fe.nf_type[f] == typ && fe.nf_anchor[f] == anc) node_face[a] = f;
  }
  int cnt = 0;
  for (int f = 0; f < fe.nnf; ++f) {
    fe.own_ptr[f] = cnt;
    for (int c = 0; c < ncomp; ++c)
      for (int a = 0; a < fe.nnode; ++a)
        if (node_face[a] == f) fe.own_lst[cnt++] = c * fe.nnode + a;
  }
  fe.own_ptr[fe.nnf] = cnt;
  // quadrature of degree 2*order, first index fastest
  fe.nq1 = order + 1;
  double x1[8], w1[8];
  gauss_legendre(fe.nq1, x1, w1);
  fe.nq = 1;
  for (int k = 0; k < dim; ++k) fe.nq *= fe.nq1;
  for (int q = 0; q < fe.nq; ++q) {
    int r = q, id[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k) {
      id[k] = r % fe.nq1;
      r /= fe.nq1;
    }
    double w = 1.0;
    for (int k = 0; k < 3; ++k) fe.qp[k][q] = 0.0;
    for (int k = 0; k < dim; ++k) {
      fe.qp[k][q] = x1[id[k]];
      w = (k == 0) ? w1[id[0]] : w * w1[id[k]];
    }
    fe.qw[q] = w;
  }
  // affine map of an axis-aligned box cell: J = diag(h/2), det J = prod h/2,
  // grad phi = (2/h) * reference gradient; reused for every (translated) cell.
  for (int q = 0; q < fe.nq; ++q) {
    double det = 1.0;
    for (int k = 0; k < dim; ++k) det *= 0.5 * h[k];
    fe.detj[q] = std::fabs(det);
    for (int a = 0; a < fe.nnode; ++a) {
      int r = a, id[3] = {0, 0, 0};
      for (int k = 0; k < dim; ++k) {
        id[k] = r % fe.np1;
        r /= fe.np1;
      }
      double v[3], dv[3];
      for (int k = 0; k < dim; ++k) lagrange(order, id[k], fe.qp[k][q], v[k], dv[k]);
      double val = 1.0;
      for (int k = 0; k < dim; ++k) val *= v[k];
      fe.shp[a][q] = val;
      for (int i = 0; i < 3; ++i) fe.grad[a][i][q] = 0.0;
      for (int i = 0; i < dim; ++i) {
        double g = dv[i];
        for (int j = 0; j < dim; ++j)
          if (j != i) g *= v[j];
        fe.grad[a][i][q] = g * (2.0 / h[i]);
      }
    }
    for (int a = 0; a < (1 << dim); ++a) {
      double val = 1.0;
      for (int k = 0; k < dim; ++k) {
        double v, dv;
        lagrange(1, (a >> k) & 1, fe.qp[k][q], v, dv);
        val *= v;
      }
      fe.gshp[a][q] = val;
    }
  }
}

// ------------------------------------------------------- DOF numbering ---
namespace {
struct MeshIdx {
  int dim;
  int64_t n[3];
  int nvtype;
  int vtype_of_face[kMaxNf];
  int64_t voff[8], vcnt[8][3];
  int64_t num_vefs;
};

MeshIdx mesh_index(const femhip_mesh_desc_t &m, const RefFE &fe) {
  MeshIdx M{};
  M.dim = m.dim;
  for (int k = 0; k < 3; ++k) M.n[k] = k < m.dim ? m.n[k] : 1;
  // vef types: anchor-0 n-faces of dim < d in polytope order
  // (Geometry/Tools/uniform_hex_mesh_generator.f90:386-421)
  int t = -1;
  int64_t off = 0;
  for (int f = 0; f < fe.nnf; ++f) {
    if (popc(fe.nf_type[f]) >= m.dim) {
      M.vtype_of_face[f] = -1;
      continue;
    }
    if (fe.nf_anchor[f] == 0) {
      ++t;
      int64_t c = 1;
      for (int k = 0; k < m.dim; ++k) {
        M.vcnt[t][k] = M.n[k] + 1 - ((fe.nf_type[f] >> k) & 1);
        c *= M.vcnt[t][k];
      }
      M.voff[t] = off;
      off += c;
    }
    M.vtype_of_face[f] = t;
  }
  M.nvtype = t + 1;
  M.num_vefs = off;
  return M;
}
}  // namespace

void number_structured(femhip_space &s) {
  const femhip_mesh_desc_t &m = s.mesh;
  const RefFE &fe = s.fe;
  MeshIdx M = mesh_index(m, fe);
  const int nd = fe.nd, nvef = fe.nnf - 1, dim = m.dim;
  s.ncells = M.n[0] * M.n[1] * M.n[2];
  s.dofs.assign((size_t)(s.ncells * nd), 0);

  // per-cell vef ids and Dirichlet flags
  auto vef_id = [&](const int64_t ijk[3], int f) -> int64_t {
    int t = M.vtype_of_face[f];
    int64_t idx = 0, mul = 1;
    for (int k = 0; k < dim; ++k) {
      idx += mul * (ijk[k] + ((fe.nf_anchor[f] >> k) & 1));
      mul *= M.vcnt[t][k];
    }
    return M.voff[t] + idx;
  };
  auto is_dirichlet = [&](const int64_t ijk[3], int f) -> bool {
    if (m.dirichlet == FEMHIP_BC_NONE) return false;
    for (int k = 0; k < dim; ++k) {
      if ((fe.nf_type[f] >> k) & 1) continue;
      int64_t v = ijk[k] + ((fe.nf_anchor[f] >> k) & 1);
      if (m.dirichlet == FEMHIP_BC_ALL_BOUNDARY && (v == 0 || v == M.n[k])) return true;
      if (m.dirichlet == FEMHIP_BC_X0_FACE && k == 0 && v == 0) return true;
    }
    return false;
  };
  // owner[v] = (cell, local face) that first numbered vef v, -1 if none yet
  std::vector<int64_t> owner((size_t)M.num_vefs, -1);
  std::vector<int8_t> owner_face((size_t)M.num_vefs, 0);
  auto copy_from_owner = [&](int32_t *cd, int f, int64_t v) {
    const int32_t *src = s.dofs.data() + owner[v] * nd;
    int fs = owner_face[v];
    int cntf = fe.own_ptr[f + 1] - fe.own_ptr[f];
    for (int k = 0; k < cntf; ++k) cd[fe.own_lst[fe.own_ptr[f] + k]] = src[fe.own_lst[fe.own_ptr[fs] + k]];
  };

  // sweep 1: strong Dirichlet DOFs get -1, -2, ... on first touch
  // (Integration/FESpace/sbm_serial_fe_space.i90:525-715)
  int32_t nfix = 0;
  int64_t ijk[3] = {0, 0, 0};
  for (int64_t c = 0; c < s.ncells; ++c) {
    int64_t r = c;
    for (int k = 0; k < dim; ++k) {
      ijk[k] = r % M.n[k];
      r /= M.n[k];
    }
    int32_t *cd = s.dofs.data() + c * nd;
    for (int f = 0; f < nvef; ++f) {
      if (!is_dirichlet(ijk, f)) continue;
      int64_t v = vef_id(ijk, f);
      if (owner[v] >= 0) {
        copy_from_owner(cd, f, v);
        continue;
      }
      int32_t before = nfix;
      for (int k = fe.own_ptr[f]; k < fe.own_ptr[f + 1]; ++k)
        if (cd[fe.own_lst[k]] >= 0) cd[fe.own_lst[k]] = -(++nfix);
      if (nfix > before) {
        owner[v] = c;
        owner_face[v] = (int8_t)f;
      }
    }
  }
  s.n_fixed = nfix;

  // sweep 2: free DOFs 1..n, cell interior first, then vefs
  // (Integration/FESpace/sbm_serial_fe_space.i90:2623-2922)
  std::fill(owner.begin(), owner.end(), -1);
  int32_t cur = 0;
  const int fc = fe.nnf - 1;
  for (int64_t c = 0; c < s.ncells; ++c) {
    int64_t r = c;
    for (int k = 0; k < dim; ++k) {
      ijk[k] = r % M.n[k];
      r /= M.n[k];
    }
    int32_t *cd = s.dofs.data() + c * nd;
    for (int k = fe.own_ptr[fc]; k < fe.own_ptr[fc + 1]; ++k) cd[fe.own_lst[k]] = ++cur;
    for (int f = 0; f < nvef; ++f) {
      int64_t v = vef_id(ijk, f);
      if (owner[v] >= 0) {
        copy_from_owner(cd, f, v);
        continue;
      }
      int32_t before = cur;
      for (int k = fe.own_ptr[f]; k < fe.own_ptr[f + 1]; ++k)
        if (cd[fe.own_lst[k]] >= 0) cd[fe.own_lst[k]] = ++cur;
      if (cur > before) {
        owner[v] = c;
        owner_face[v] = (int8_t)f;
      }
    }
  }
  s.n_free = cur;
}

// ------------------------------------------------------------ CSR pattern ---
// The pattern sort_and_compress yields is the sorted set of distinct (row, col)
// pairs of free DOFs sharing a cell (Dirichlet columns/rows dropped,
// base_sparse_matrix.f90:3499-3510); build it row by row from the DOF -> cell
// adjacency, in parallel over rows.
void build_pattern(femhip_operator &op) {
  femhip_space &s = *op.space;
  const int nd = s.fe.nd;
  const int32_t n = s.n_free;
  const int64_t nc = s.ncells;
  std::vector<int64_t> aptr((size_t)n + 1, 0);
  for (int64_t i = 0; i < nc * nd; ++i) {
    int32_t g = s.dofs[(size_t)i];
    if (g > 0) aptr[(size_t)g]++;
  }
  for (int32_t r = 0; r < n; ++r) aptr[(size_t)r + 1] += aptr[(size_t)r];
  std::vector<int64_t> fill(aptr.begin(), aptr.end() - 1);
  std::vector<int64_t> acell((size_t)aptr[(size_t)n]);
  for (int64_t c = 0; c < nc; ++c)
    for (int l = 0; l < nd; ++l) {
      int32_t g = s.dofs[(size_t)(c * nd + l)];
      if (g > 0) acell[(size_t)fill[(size_t)g - 1]++] = c;
    }
  fill.clear();
  fill.shrink_to_fit();

  unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<int64_t> rowcnt((size_t)n + 1, 0);
  auto row_cols = [&](int32_t r, std::vector<int32_t> &buf) {
    buf.clear();
    for (int64_t k = aptr[(size_t)r]; k < aptr[(size_t)r + 1]; ++k) {
      const int32_t *cd = s.dofs.data() + acell[(size_t)k] * nd;
      for (int l = 0; l < nd; ++l)
        if (cd[l] > 0) buf.push_back(cd[l] - 1);
    }
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
  auto par = [&](auto &&body) {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; ++t)
      th.emplace_back([&, t] {
        std::vector<int32_t> buf;
        int64_t lo = (int64_t)n * t / nth, hi = (int64_t)n * (t + 1) / nth;
        for (int64_t r = lo; r < hi; ++r) body((int32_t)r, buf);
      });
    for (auto &x : th) x.join();
  };
  par([&](int32_t r, std::vector<int32_t> &buf) {
    row_cols(r, buf);
    rowcnt[(size_t)r + 1] = (int64_t)buf.size();
  });
  for (int32_t r = 0; r < n; ++r) rowcnt[(size_t)r + 1] += rowcnt[(size_t)r];
  FH_CHECK(rowcnt[(size_t)n] < (int64_t)INT32_MAX, FEMHIP_ERR_INVALID,
           "CSR pattern exceeds 2^31-1 entries (32-bit irp/ja of csr_sparse_matrix_t)");
  op.n = n;
  op.nnz_full = rowcnt[(size_t)n];
  op.h_rowptr.resize((size_t)n + 1);
  op.h_col.resize((size_t)op.nnz_full);
  for (int32_t r = 0; r <= n; ++r) op.h_rowptr[(size_t)r] = (int32_t)rowcnt[(size_t)r];
  par([&](int32_t r, std::vector<int32_t> &buf) {
    row_cols(r, buf);
    std::copy(buf.begin(), buf.end(), op.h_col.begin() + rowcnt[(size_t)r]);
  });
  // upper-triangle view for symmetric storage (entries with col >= row)
  op.h_upper_map.clear();
  if (op.symmetric_storage) {
    op.h_upper_map.reserve((size_t)((op.nnz_full + n) / 2));
    for (int32_t r = 0; r < n; ++r)
      for (int64_t k = op.h_rowptr[(size_t)r]; k < op.h_rowptr[(size_t)r + 1]; ++k)
        if (op.h_col[(size_t)k] >= r) op.h_upper_map.push_back(k);
  }
  op.nnz_upper = (int64_t)op.h_upper_map.size();
}

}  // namespace femhip
