// space.cpp -- reference FE tables of the FE space (host): polytope n-faces
// and their own DOFs, quadrature, Lagrange shape functions, affine map of the
// first cell.  The DOF numbering and the CSR pattern are built on the device
// (setup.hip).
#include <algorithm>
#include <cmath>
#include <cstring>

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

void build_reffe(RefFE &fe, int dim, int order, int ncomp, const double h[3], int nq1) {
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
  // quadrature of degree 2*order (or nq1 points per direction when a
  // multi-field cell integrates every field with its highest order's rule),
  // first index fastest
  fe.nq1 = nq1 > order ? nq1 : order + 1;
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
      for (int i = 0; i < 3; ++i) fe.grad[a][i][q] = fe.dshp[a][i][q] = 0.0;
      for (int i = 0; i < dim; ++i) {
        double g = dv[i];
        for (int j = 0; j < dim; ++j)
          if (j != i) g *= v[j];
        fe.dshp[a][i][q] = g;
        fe.grad[a][i][q] = g * (2.0 / h[i]);
      }
    }
    for (int a = 0; a < (1 << dim); ++a) {
      double v[3], dv[3], val = 1.0;
      for (int k = 0; k < dim; ++k) {
        lagrange(1, (a >> k) & 1, fe.qp[k][q], v[k], dv[k]);
        val *= v[k];
      }
      fe.gshp[a][q] = val;
      for (int i = 0; i < 3; ++i) fe.gdshp[a][i][q] = 0.0;
      for (int i = 0; i < dim; ++i) {
        double g = dv[i];
        for (int j = 0; j < dim; ++j)
          if (j != i) g *= v[j];
        fe.gdshp[a][i][q] = g;
      }
    }
  }
}

}  // namespace femhip
