/*
 * fempar_oracle.c -- CPU restatement of FEMPAR's one-subdomain FE hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see fempar_oracle.h): the checker for the HIP
 * product path, never the product.  Plain C99, single thread, reference loop
 * order throughout.  File:line citations are relative to the reference's
 * Sources/Lib unless they start with Sources/.
 */
#include "fempar_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXNF 27          /* n-faces of a hexahedron incl. the cell itself */
#define MAXND 81          /* Q2 scalar (27) or vector Q1 3D (24); Q2 vector 3D = 81 */
#define MAXNQ 27

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n); abort(); }
  return p;
}
static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz ? sz : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}
void orc_free(void *p) { free(p); }

static int popcount(int v) { int c = 0; while (v) { c += v & 1; v >>= 1; } return c; }

struct orc_space {
  orc_problem_t p;
  int d;
  /* polytope: n-faces sorted by (dimension, bit value) -- sbm_polytope.i90:175-227 */
  int nnf;                 /* number of n-faces incl. the cell */
  int nf_type[MAXNF];      /* free-direction mask */
  int nf_anchor[MAXNF];    /* anchor-vertex bits */
  int nf_vtype[MAXNF];     /* mesh vef type index (uniform_hex_mesh_generator.f90:386-391) */
  int ntypes;
  int64_t type_off[8];
  int64_t type_cnt[8][3];
  int64_t num_vefs;
  /* reference FE (lagrangian_reference_fe_fill_scalar, sbm_lagrangian_reference_fe.i90:1549-1760) */
  int order, np1, nnode, ncomp, nd;
  int own_ptr[MAXNF + 1];
  int own_lst[MAXND];
  /* quadrature + tables */
  int nq1, nq;
  double qp[3][MAXNQ];     /* quadrature coordinates (dim x nq) */
  double qw[MAXNQ];
  double shp[MAXND][MAXNQ];          /* scalar shape values N_a(q) */
  double dshp[MAXND][3][MAXNQ];      /* reference gradients dN_a/dxi_j (q) */
  double gshp[8][MAXNQ];             /* geometric Q1 shape values */
  double gdshp[8][3][MAXNQ];         /* geometric Q1 gradients */
  /* affine map of the first cell, reused by translation (sbm_fe_cell_iterator.i90:970-987) */
  double detj[MAXNQ];
  double grad[MAXND][3][MAXNQ];      /* physical scalar gradients (sbm_lagrangian_reference_fe.i90:1334-1341) */
  /* space */
  int64_t ncells;
  int32_t *dofs;           /* ncells x nd, signed 1-based */
  int nfree, nfixed;
  int nq1_req;             /* quadrature points per direction requested by a multi-field space (0: order + 1) */
};

/* ------------------------------------------------------------------ mesh --- */

static void cell_ijk(const orc_space_t *s, int64_t c, int ijk[3]) {
  /* spatial_to_ijk_numbering, x fastest (uniform_hex_mesh_generator.f90:801-815) */
  int64_t j = c;
  ijk[0] = ijk[1] = ijk[2] = 0;
  for (int d = 0; d < s->d; ++d) { ijk[d] = (int)(j % s->p.n[d]); j /= s->p.n[d]; }
}

static double vertex_coord(const orc_space_t *s, int d, int i) {
  /* coordinates = i/n then mapped to the domain limits (uniform_hex_mesh_generator.f90:608-612, 633-635) */
  double c = (double)i / (double)s->p.n[d];
  return (s->p.lim[2 * d + 1] - s->p.lim[2 * d]) * c + s->p.lim[2 * d];
}

static int64_t vef_gid(const orc_space_t *s, const int ijk[3], int f) {
  /* lst_vefs_lids (uniform_hex_mesh_generator.f90:441-456), 0-based */
  int t = s->nf_vtype[f];
  int64_t idx = 0, mul = 1;
  for (int d = 0; d < s->d; ++d) {
    int v = ijk[d] + ((s->nf_anchor[f] >> d) & 1);
    idx += mul * v;
    mul *= s->type_cnt[t][d];
  }
  return s->type_off[t] + idx;
}

static int vef_set_id(const orc_space_t *s, const int ijk[3], int f) {
  /* boundary detection (sbm_fine_triangulation.i90:453-527): a vef is at the boundary
     when it lies on a facet with one cell around, i.e. a fixed direction sits at 0 or n.
     test_poisson_driver.f90:193-204 then sets set_id = 1 on boundary vefs. */
  if (s->p.bc_kind == ORC_BC_NONE) return 0;
  for (int d = 0; d < s->d; ++d) {
    if ((s->nf_type[f] >> d) & 1) continue;           /* free direction */
    int v = ijk[d] + ((s->nf_anchor[f] >> d) & 1);
    if (s->p.bc_kind == ORC_BC_ALL_BOUNDARY) {
      if (v == 0 || v == s->p.n[d]) return 1;
    } else if (s->p.bc_kind == ORC_BC_X0_FACE) {
      if (d == 0 && v == 0) return 1;
    }
  }
  return 0;
}

static void build_polytope(orc_space_t *s) {
  int d = s->d, nn = 0;
  int val[MAXNF], typ[MAXNF], anc[MAXNF];
  for (int t = 0; t < (1 << d); ++t)
    for (int a = 0; a < (1 << d); ++a)
      if ((a & t) == 0) { typ[nn] = t; anc[nn] = a; val[nn] = (t << d) | a; ++nn; }
  /* sort_n_faces: by n-face dimension, then by bit-based index (sbm_polytope.i90:206-217) */
  for (int i = 1; i < nn; ++i) {
    int k = i;
    while (k > 0) {
      int dk = popcount(typ[k]), dp = popcount(typ[k - 1]);
      if (dk < dp || (dk == dp && val[k] < val[k - 1])) {
        int tv = val[k]; val[k] = val[k - 1]; val[k - 1] = tv;
        int tt = typ[k]; typ[k] = typ[k - 1]; typ[k - 1] = tt;
        int ta = anc[k]; anc[k] = anc[k - 1]; anc[k - 1] = ta;
        --k;
      } else break;
    }
  }
  s->nnf = nn;
  for (int i = 0; i < nn; ++i) { s->nf_type[i] = typ[i]; s->nf_anchor[i] = anc[i]; }
  /* vef types: n-faces of dim < d with anchor 0, in polytope order
     (uniform_hex_mesh_generator.f90:386-421) */
  int it = -1;
  int64_t off = 0;
  for (int i = 0; i < nn; ++i) {
    if (popcount(typ[i]) >= d) { s->nf_vtype[i] = -1; continue; }
    if (anc[i] == 0) {
      ++it;
      int64_t cnt = 1;
      for (int k = 0; k < d; ++k) {
        s->type_cnt[it][k] = s->p.n[k] + 1 - ((typ[i] >> k) & 1);
        cnt *= s->type_cnt[it][k];
      }
      s->type_off[it] = off;
      off += cnt;
    }
    s->nf_vtype[i] = it;
  }
  s->ntypes = it + 1;
  s->num_vefs = off;
}

/* ---------------------------------------------------- reference FE tables --- */

void orc_gauss_legendre(int n, double *x, double *w) {
  /* Gauss-Legendre abscissae by Newton iteration on the three-term recurrence,
     ascending; weights 2/((1-x)(1+x)P_n'(x)^2) renormalised to sum 2 exactly as
     legendre_compute_glr does (sbm_quadrature.i90:857-862).  The reference reaches
     the same roots with the Glaser-Liu-Rokhlin ODE method; the two agree to
     rounding (pinned against test_quadratures.f90 tables). */
  const double pi = 3.14159265358979323846;
  double *dp = (double *)xmalloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    double z = -cos(pi * (i + 0.75) / (n + 0.5));
    double pp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      if (n == 0) p1 = 1.0;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1; p1 = p2;
      }
      double pn = (n == 1) ? z : p1, pnm1 = (n == 1) ? 1.0 : p0;
      pp = n * (z * pn - pnm1) / (z * z - 1.0);
      double dz = pn / pp;
      z -= dz;
      if (fabs(dz) < 1e-17) break;
    }
    {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1; p1 = p2;
      }
      double pn = (n == 1) ? z : p1, pnm1 = (n == 1) ? 1.0 : p0;
      pp = n * (z * pn - pnm1) / (z * z - 1.0);
    }
    if (n % 2 == 1 && i == n / 2) z = 0.0;
    x[i] = z;
    dp[i] = pp;
  }
  /* symmetrise: exact antisymmetry of the rule */
  for (int i = 0; i < n / 2; ++i) {
    double a = 0.5 * (x[n - 1 - i] - x[i]);
    x[i] = -a; x[n - 1 - i] = a;
    double b = 0.5 * (fabs(dp[i]) + fabs(dp[n - 1 - i]));
    dp[i] = b; dp[n - 1 - i] = b;
  }
  double sum = 0.0;
  for (int i = 0; i < n; ++i) {
    w[i] = 2.0 / (1.0 - x[i]) / (1.0 + x[i]) / dp[i] / dp[i];
    sum += w[i];
  }
  for (int i = 0; i < n; ++i) w[i] = 2.0 * w[i] / sum;
  free(dp);
}

static void lagrange_1d(int order, int i, double x, double *v, double *dv) {
  /* lagrange_polynomial_generate_basis + get_values (sbm_polynomial.i90:61-111):
     l_i(x) = c_i * prod_{k != i} (x - t_k), t equispaced on [-1,1], with the
     value/derivative recurrence of get_values. */
  double t[8];
  for (int k = 0; k <= order; ++k) t[k] = (double)k;
  if (order > 0) for (int k = 0; k <= order; ++k) t[k] = (2.0 / order) * t[k] - 1.0;
  double coef[8];
  int m = 0;
  for (int k = 0; k <= order; ++k) if (k != i) coef[m++] = t[k];
  double c = 1.0;
  for (int k = 0; k <= order; ++k) if (k != i) c = c * (t[i] - t[k]);
  c = 1.0 / c;
  double px[3] = {1.0, 0.0, 0.0};
  for (int k = 0; k < order; ++k) {
    for (int j = 2; j >= 1; --j) px[j] = px[j] * (x - coef[k]) + px[j - 1];
    px[0] = px[0] * (x - coef[k]);
  }
  *v = px[0] * c;
  *dv = px[1] * c;
}

static void build_reffe(orc_space_t *s) {
  int d = s->d;
  s->np1 = s->order + 1;
  s->nnode = 1;
  for (int k = 0; k < d; ++k) s->nnode *= s->np1;
  s->nd = s->nnode * s->ncomp;
  /* own DOFs per n-face: the n-face of node m is (free dirs = interior coords,
     anchor = coords at the upper end); own nodes in node order; vector fields
     extend component-major (lagrangian_reference_fe_extend_list_components,
     sbm_lagrangian_reference_fe.i90:1250-1286). */
  int node_nf[MAXND];
  for (int a = 0; a < s->nnode; ++a) {
    int rem = a, typ = 0, anc = 0;
    for (int k = 0; k < d; ++k) {
      int mk = rem % s->np1; rem /= s->np1;
      if (mk > 0 && mk < s->order) typ |= 1 << k;
      else if (mk == s->order) anc |= 1 << k;
    }
    node_nf[a] = -1;
    for (int f = 0; f < s->nnf; ++f)
      if (s->nf_type[f] == typ && s->nf_anchor[f] == anc) node_nf[a] = f;
  }
  int cnt = 0;
  for (int f = 0; f < s->nnf; ++f) {
    s->own_ptr[f] = cnt;
    for (int c = 0; c < s->ncomp; ++c)
      for (int a = 0; a < s->nnode; ++a)
        if (node_nf[a] == f) s->own_lst[cnt++] = c * s->nnode + a;
  }
  s->own_ptr[s->nnf] = cnt;

  /* quadrature: degree 2*order -> order+1 points per direction
     (sbm_lagrangian_reference_fe.i90:1414-1424, sbm_hex_lagrangian_reference_fe.i90:320-328),
     first index fastest (sbm_quadrature.i90:797-815) */
  s->nq1 = s->nq1_req > s->order ? s->nq1_req : s->order + 1;
  double x1[8], w1[8];
  orc_gauss_legendre(s->nq1, x1, w1);
  s->nq = 1;
  for (int k = 0; k < d; ++k) s->nq *= s->nq1;
  for (int q = 0; q < s->nq; ++q) {
    int rem = q;
    double w = 1.0;
    for (int k = 0; k < 3; ++k) s->qp[k][q] = 0.0;
    int id[3] = {0, 0, 0};
    for (int k = 0; k < d; ++k) { id[k] = rem % s->nq1; rem /= s->nq1; }
    for (int k = 0; k < d; ++k) { s->qp[k][q] = x1[id[k]]; w = (k == 0) ? w1[id[0]] : w * w1[id[k]]; }
    s->qw[q] = w;
  }
  /* tensor-product values/gradients (sbm_tensor_product_polynomial.i90:63-127) */
  for (int q = 0; q < s->nq; ++q) {
    for (int a = 0; a < s->nnode; ++a) {
      int rem = a, id[3] = {0, 0, 0};
      for (int k = 0; k < d; ++k) { id[k] = rem % s->np1; rem /= s->np1; }
      double v[3], dv[3];
      for (int k = 0; k < d; ++k) lagrange_1d(s->order, id[k], s->qp[k][q], &v[k], &dv[k]);
      double val = 1.0;
      for (int k = 0; k < d; ++k) val = val * v[k];
      s->shp[a][q] = val;
      for (int i = 0; i < 3; ++i) s->dshp[a][i][q] = 0.0;
      for (int i = 0; i < d; ++i) {
        double g = dv[i];
        for (int j = 0; j < d; ++j) if (j != i) g = g * v[j];
        s->dshp[a][i][q] = g;
      }
    }
    for (int a = 0; a < (1 << d); ++a) {
      double v[3], dv[3];
      for (int k = 0; k < d; ++k) lagrange_1d(1, (a >> k) & 1, s->qp[k][q], &v[k], &dv[k]);
      double val = 1.0;
      for (int k = 0; k < d; ++k) val = val * v[k];
      s->gshp[a][q] = val;
      for (int i = 0; i < 3; ++i) s->gdshp[a][i][q] = 0.0;
      for (int i = 0; i < d; ++i) {
        double g = dv[i];
        for (int j = 0; j < d; ++j) if (j != i) g = g * v[j];
        s->gdshp[a][i][q] = g;
      }
    }
  }
}

static void cell_vertices(const orc_space_t *s, int64_t c, double xv[8][3]) {
  /* fe_cell_iterator_t%get_nodes_coordinates (sbm_fe_cell_iterator.i90:966-967):
     the triangulation's vertex coordinates, generated (uniform) or given */
  int ijk[3];
  cell_ijk(s, c, ijk);
  for (int a = 0; a < (1 << s->d); ++a) {
    for (int k = 0; k < 3; ++k) xv[a][k] = 0.0;
    if (s->p.vcoords) {
      int64_t v = 0, mul = 1;
      for (int k = 0; k < s->d; ++k) {
        v += mul * (ijk[k] + ((a >> k) & 1));
        mul *= s->p.n[k] + 1;
      }
      for (int k = 0; k < s->d; ++k) xv[a][k] = s->p.vcoords[v * s->d + k];
    } else {
      for (int k = 0; k < s->d; ++k) xv[a][k] = vertex_coord(s, k, ijk[k] + ((a >> k) & 1));
    }
  }
}

/* cell_map_update (sbm_cell_map.i90:245-369) on cell c: Jacobian as the linear
   combination of geometric gradients and node coordinates (field.f90:763-787),
   inverse and determinant by invmtx (sbm_cell_integrator.i90:448-560), physical
   gradients (sbm_lagrangian_reference_fe.i90:1334-1341). */
static void cell_map(const orc_space_t *s, int64_t c, double detj[MAXNQ], double grad[MAXND][3][MAXNQ]) {
  double xv[8][3];
  cell_vertices(s, c, xv);
  int d = s->d;
  for (int q = 0; q < s->nq; ++q) {
    double J[3][3] = {{0}};
    for (int j = 0; j < 3; ++j) {
      for (int l = 0; l < 3; ++l) J[l][j] = 0.0;
      for (int k = 0; k < (1 << d); ++k) {
        double alpha = s->gdshp[k][j][q];
        for (int l = 0; l < 3; ++l) J[l][j] = J[l][j] + alpha * xv[k][l];
      }
    }
    double B[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double det;
    if (d == 2) {
      det = J[0][0] * J[1][1] - J[1][0] * J[0][1];
      double den = 1.0 / det;
      B[0][0] = J[1][1] * den; B[1][1] = J[0][0] * den;
      B[1][0] = -J[1][0] * den; B[0][1] = -J[0][1] * den;
    } else {
      double t1 = J[1][1] * J[2][2] - J[2][1] * J[1][2];
      double t2 = -J[1][0] * J[2][2] + J[2][0] * J[1][2];
      double t3 = J[1][0] * J[2][1] - J[2][0] * J[1][1];
      det = J[0][0] * t1 + J[0][1] * t2 + J[0][2] * t3;
      double den = 1.0 / det;
      B[0][0] = t1 * den; B[1][0] = t2 * den; B[2][0] = t3 * den;
      B[1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) * den;
      B[2][1] = (-J[0][0] * J[2][1] + J[0][1] * J[2][0]) * den;
      B[2][2] = (J[0][0] * J[1][1] - J[1][0] * J[0][1]) * den;
      B[0][1] = (-J[0][1] * J[2][2] + J[2][1] * J[0][2]) * den;
      B[0][2] = (J[0][1] * J[1][2] - J[1][1] * J[0][2]) * den;
      B[1][2] = (-J[0][0] * J[1][2] + J[1][0] * J[0][2]) * den;
    }
    detj[q] = fabs(det);
    /* phy(i) += invJ(j,i) * ref(j)  (sbm_lagrangian_reference_fe.i90:1334-1341) */
    for (int a = 0; a < s->nnode; ++a)
      for (int i = 0; i < 3; ++i) {
        double g = 0.0;
        for (int j = 0; j < 3; ++j) g = g + B[j][i] * s->dshp[a][j][q];
        grad[a][i][q] = g;
      }
  }
}

static void build_map(orc_space_t *s) {
  /* the first cell's map; on the generator's uniform mesh every later cell is
     "only translated" and reuses it (sbm_fe_cell_iterator.i90:970-987) */
  cell_map(s, 0, s->detj, s->grad);
}

/* ------------------------------------------------------- DOF numbering --- */

static void number_dofs(orc_space_t *s) {
  int nd = s->nd, nvef = s->nnf - 1;
  s->dofs = (int32_t *)xcalloc((size_t)(s->ncells * nd), sizeof(int32_t));
  int64_t *vis_cell = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)s->num_vefs);
  int8_t *vis_lid = (int8_t *)xmalloc((size_t)s->num_vefs);
  for (int64_t g = 0; g < s->num_vefs; ++g) vis_cell[g] = -1;

  /* strong Dirichlet sweep (sbm_serial_fe_space.i90:525-715): cells in order,
     vefs in polytope order, negative ids on first visit, fetched afterwards. */
  int ndir = 0;
  int ijk[3];
  for (int64_t c = 0; c < s->ncells; ++c) {
    cell_ijk(s, c, ijk);
    int32_t *cd = s->dofs + c * nd;
    for (int f = 0; f < nvef; ++f) {
      if (vef_set_id(s, ijk, f) == 0) continue;
      int64_t g = vef_gid(s, ijk, f);
      int prev = ndir;
      if (vis_cell[g] == -1) {
        /* generate_own_dofs_vef_component_wise: all components coded as fixed */
        for (int k = s->own_ptr[f]; k < s->own_ptr[f + 1]; ++k) {
          int l = s->own_lst[k];
          if (cd[l] >= 0) cd[l] = -(++ndir);
        }
      } else {
        /* fetch_own_dofs_vef_from_source_fe: one node per n-face for Q1/Q2 */
        const int32_t *src = s->dofs + vis_cell[g] * nd;
        int fs = vis_lid[g];
        for (int k = 0; k < s->own_ptr[f + 1] - s->own_ptr[f]; ++k)
          cd[s->own_lst[s->own_ptr[f] + k]] = src[s->own_lst[s->own_ptr[fs] + k]];
      }
      if (vis_cell[g] == -1 && prev < ndir) { vis_cell[g] = c; vis_lid[g] = (int8_t)f; }
    }
  }
  s->nfixed = ndir;

  /* free DOFs (serial_fe_space_list_dofs, sbm_serial_fe_space.i90:2623-2922): cell
     interior first, then vefs; first visit numbers non-fixed own DOFs. */
  for (int64_t g = 0; g < s->num_vefs; ++g) vis_cell[g] = -1;
  int cur = 0;
  for (int64_t c = 0; c < s->ncells; ++c) {
    cell_ijk(s, c, ijk);
    int32_t *cd = s->dofs + c * nd;
    int fc = s->nnf - 1;
    for (int k = s->own_ptr[fc]; k < s->own_ptr[fc + 1]; ++k) cd[s->own_lst[k]] = ++cur;
    for (int f = 0; f < nvef; ++f) {
      int64_t g = vef_gid(s, ijk, f);
      if (vis_cell[g] == -1) {
        int prev = cur;
        for (int k = s->own_ptr[f]; k < s->own_ptr[f + 1]; ++k) {
          int l = s->own_lst[k];
          if (cd[l] >= 0) cd[l] = ++cur;
        }
        if (prev < cur) { vis_cell[g] = c; vis_lid[g] = (int8_t)f; }
      } else {
        const int32_t *src = s->dofs + vis_cell[g] * nd;
        int fs = vis_lid[g];
        for (int k = 0; k < s->own_ptr[f + 1] - s->own_ptr[f]; ++k)
          cd[s->own_lst[s->own_ptr[f] + k]] = src[s->own_lst[s->own_ptr[fs] + k]];
      }
    }
  }
  s->nfree = cur;
  free(vis_cell);
  free(vis_lid);
}

static orc_space_t *space_create(const orc_problem_t *prob, int nq1_req);

orc_space_t *orc_space_create(const orc_problem_t *prob) { return space_create(prob, 0); }

static orc_space_t *space_create(const orc_problem_t *prob, int nq1_req) {
  if (prob->dim != 2 && prob->dim != 3) return NULL;
  if (prob->order < 1 || prob->order > 2) return NULL;
  if (prob->ncomp != 1 && prob->ncomp != prob->dim) return NULL;
  orc_space_t *s = (orc_space_t *)xcalloc(1, sizeof(orc_space_t));
  s->p = *prob;
  s->d = prob->dim;
  if (s->d == 2) { s->p.n[2] = 1; s->p.lim[4] = 0.0; s->p.lim[5] = 0.0; }
  s->order = prob->order;
  s->ncomp = prob->ncomp;
  s->nq1_req = nq1_req;
  s->ncells = 1;
  for (int k = 0; k < s->d; ++k) s->ncells *= prob->n[k];
  build_polytope(s);
  build_reffe(s);
  build_map(s);
  number_dofs(s);
  return s;
}

void orc_space_free(orc_space_t *s) {
  if (!s) return;
  free(s->dofs);
  free(s);
}

void orc_space_sizes(const orc_space_t *s, int64_t *num_cells, int *nd, int *nq, int *n_free,
                     int *n_fixed) {
  *num_cells = s->ncells; *nd = s->nd; *nq = s->nq; *n_free = s->nfree; *n_fixed = s->nfixed;
}

void orc_space_get_dofs(const orc_space_t *s, int32_t *out) {
  memcpy(out, s->dofs, sizeof(int32_t) * (size_t)(s->ncells * s->nd));
}

void orc_reference_tables(const orc_space_t *s, double *qpts, double *qw, double *shape,
                          double *dshape) {
  for (int q = 0; q < s->nq; ++q) {
    for (int k = 0; k < s->d; ++k) qpts[q * s->d + k] = s->qp[k][q];
    qw[q] = s->qw[q];
    for (int a = 0; a < s->nnode; ++a) {
      shape[q * s->nnode + a] = s->shp[a][q];
      for (int k = 0; k < s->d; ++k) dshape[(q * s->nnode + a) * s->d + k] = s->dshp[a][k][q];
    }
  }
}

/* --------------------------------------------------------- physics ------ */

static const double PI = 3.14159265358979323846;

static double k_value(const orc_physics_t *ph, int d, const double x[3], int64_t cell, int q, int nq) {
  switch (ph->k_kind) {
    case ORC_K_SIN3: {
      double v = 1.0 + 0.5 * sin(2 * PI * x[0]) * sin(2 * PI * x[1]) * (d == 3 ? sin(2 * PI * x[2]) : 1.0);
      return v;
    }
    case ORC_K_CELL: return ph->k_cell[cell];
    case ORC_K_QPOINT: return ph->k_qp[cell * nq + q];
    default: return ph->k0;
  }
}

static double f_value(const orc_physics_t *ph, int d, const double x[3], int64_t cell, int q, int nq) {
  switch (ph->f_kind) {
    case ORC_F_MANUFACTURED: {
      /* f = -div(k grad u) = k * d pi^2 u - grad k . grad u, u = prod sin(pi x_k) */
      double su[3], cu[3];
      for (int k = 0; k < 3; ++k) { su[k] = (k < d) ? sin(PI * x[k]) : 1.0; cu[k] = (k < d) ? cos(PI * x[k]) : 0.0; }
      double u = su[0] * su[1] * su[2];
      double kk = k_value(ph, d, x, cell, q, nq);
      double f = kk * d * PI * PI * u;
      if (ph->k_kind == ORC_K_SIN3) {
        double s2[3], c2[3];
        for (int k = 0; k < 3; ++k) { s2[k] = (k < d) ? sin(2 * PI * x[k]) : 1.0; c2[k] = (k < d) ? cos(2 * PI * x[k]) : 0.0; }
        double gk[3], gu[3];
        gk[0] = PI * c2[0] * s2[1] * s2[2];
        gk[1] = PI * s2[0] * c2[1] * s2[2];
        gk[2] = PI * s2[0] * s2[1] * c2[2];
        gu[0] = PI * cu[0] * su[1] * su[2];
        gu[1] = PI * su[0] * cu[1] * su[2];
        gu[2] = PI * su[0] * su[1] * cu[2];
        double dot = 0.0;
        for (int k = 0; k < d; ++k) dot += gk[k] * gu[k];
        f -= dot;
      }
      return f;
    }
    case ORC_F_CELL_SIN: {
      double u = 1.0;
      for (int k = 0; k < d; ++k) u *= sin(PI * x[k]);
      return ph->f0 * k_value(ph, d, x, cell, q, nq) * d * PI * PI * u;
    }
    case ORC_F_QPOINT: return ph->f_qp[cell * nq + q];
    default: return ph->f0;
  }
}

static double g_value(const orc_physics_t *ph, int d, const double x[3], int comp) {
  (void)comp;
  if (ph->g_kind == ORC_G_LINEAR) {
    double v = 0.0;
    for (int k = 0; k < d; ++k) v += x[k];
    return v;
  }
  return 0.0;
}

static void qpoint_coords(const orc_space_t *s, int64_t c, double xq[MAXNQ][3]) {
  /* cell_map_compute_quadrature_points_coordinates (sbm_cell_map.i90:673-687) */
  double xv[8][3];
  cell_vertices(s, c, xv);
  for (int q = 0; q < s->nq; ++q) {
    for (int k = 0; k < 3; ++k) xq[q][k] = 0.0;
    for (int a = 0; a < (1 << s->d); ++a)
      for (int k = 0; k < 3; ++k) xq[q][k] = xq[q][k] + xv[a][k] * s->gshp[a][q];
  }
}

void orc_quadrature_points(const orc_space_t *s, double *out) {
  double xq[MAXNQ][3];
  for (int64_t c = 0; c < s->ncells; ++c) {
    qpoint_coords(s, c, xq);
    for (int q = 0; q < s->nq; ++q)
      for (int k = 0; k < s->d; ++k) out[(c * s->nq + q) * s->d + k] = xq[q][k];
  }
}

/* the cell's map: the space's first-cell map (uniform mesh, every cell a
   translate) or the cell's own (given vertex coordinates) */
typedef struct { double detj[MAXNQ]; double grad[MAXND][3][MAXNQ]; } cell_geo_t;
static const cell_geo_t *cell_geometry(const orc_space_t *s, int64_t c, cell_geo_t *buf) {
  if (!s->p.vcoords) return (const cell_geo_t *)0;
  cell_map(s, c, buf->detj, buf->grad);
  return buf;
}

void orc_element(const orc_space_t *s, const orc_physics_t *ph, int64_t c, double *elmat,
                 double *elvec) {
  int nd = s->nd, nn = s->nnode, d = s->d;
  double xq[MAXNQ][3];
  qpoint_coords(s, c, xq);
  cell_geo_t geo;
  const cell_geo_t *g = cell_geometry(s, c, &geo);
  const double *detj = g ? g->detj : s->detj;
  const double (*grad)[3][MAXNQ] = g ? g->grad : s->grad;
  for (int i = 0; i < nd * nd; ++i) elmat[i] = 0.0;
  for (int i = 0; i < nd; ++i) elvec[i] = 0.0;
  if (ph->kind == ORC_PHYS_POISSON) {
    /* poisson_discrete_integration.f90:116-130 (q outer, i, j), with the
       coefficient at the quadrature point as in the heterogeneous-coefficient
       plugin of par_test_pb_bddc_poisson */
    for (int q = 0; q < s->nq; ++q) {
      double factor = detj[q] * s->qw[q];
      double kq = k_value(ph, d, xq[q], c, q, s->nq);
      for (int i = 0; i < nd; ++i)
        for (int j = 0; j < nd; ++j) {
          double dot = 0.0;
          for (int k = 0; k < 3; ++k) dot += grad[j][k][q] * grad[i][k][q];
          elmat[i + j * nd] = elmat[i + j * nd] + factor * kq * dot;
        }
      double fq = f_value(ph, d, xq[q], c, q, s->nq);
      for (int i = 0; i < nd; ++i) elvec[i] = elvec[i] + factor * fq * s->shp[i][q];
    }
  } else {
    /* sbm_irreducible_discrete_integration.i90:64-130: sigma = lambda tr(eps) I + 2 mu eps,
       elmat(i,j) += dV * grad(phi_i) : sigma(phi_j); elvec(i) += dV * f . phi_i */
    double E = ph->E, nu = ph->nu;
    double lambda = (nu * E) / ((1 + nu) * (1 - 2 * nu));
    double mu = E / (2 * (1 + nu));
    for (int q = 0; q < s->nq; ++q) {
      double dV = detj[q] * s->qw[q];
      for (int i = 0; i < nd; ++i) {
        int ci = i / nn, ai = i % nn;
        for (int j = 0; j < nd; ++j) {
          int cj = j / nn, aj = j % nn;
          /* grad(phi_j)(r,s) = e_cj(r) * dN_aj/dx_s ; eps = sym part */
          double G[3][3] = {{0}}, eps[3][3], sig[3][3];
          for (int k = 0; k < 3; ++k) G[cj][k] = grad[aj][k][q];
          for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) eps[a][b] = 0.5 * (G[a][b] + G[b][a]);
          double tr = eps[0][0] + eps[1][1] + eps[2][2];
          for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) sig[a][b] = lambda * tr * (a == b ? 1.0 : 0.0) + 2 * mu * eps[a][b];
          double dc = 0.0;
          for (int k = 0; k < 3; ++k) dc += grad[ai][k][q] * sig[ci][k];
          elmat[i + j * nd] = elmat[i + j * nd] + dV * dc;
        }
        elvec[i] = elvec[i] + dV * ph->body[ci] * s->shp[ai][q];
      }
    }
  }
}

void orc_space_fixed_values(const orc_space_t *s, const orc_physics_t *ph, double *out) {
  /* nodal interpolation of g at the physical node positions (cell map of the
     reference Lagrange nodes), sbm_serial_fe_space.i90:997-1176 */
  int nd = s->nd, nn = s->nnode, d = s->d;
  for (int i = 0; i < s->nfixed; ++i) out[i] = 0.0;
  for (int64_t c = 0; c < s->ncells; ++c) {
    const int32_t *cd = s->dofs + c * nd;
    int any = 0;
    for (int l = 0; l < nd; ++l) if (cd[l] < 0) { any = 1; break; }
    if (!any) continue;
    double xv[8][3];
    cell_vertices(s, c, xv);
    for (int l = 0; l < nd; ++l) {
      if (cd[l] >= 0) continue;
      int a = l % nn, comp = l / nn, rem = a;
      double xi[3] = {0, 0, 0};
      for (int k = 0; k < d; ++k) { int m = rem % s->np1; rem /= s->np1; xi[k] = (s->order > 0) ? (2.0 / s->order) * m - 1.0 : 0.0; }
      double x[3] = {0, 0, 0};
      for (int v = 0; v < (1 << d); ++v) {
        double N = 1.0;
        for (int k = 0; k < d; ++k) N *= ((v >> k) & 1) ? 0.5 * (1.0 + xi[k]) : 0.5 * (1.0 - xi[k]);
        for (int k = 0; k < 3; ++k) x[k] += xv[v][k] * N;
      }
      out[-cd[l] - 1] = g_value(ph, d, x, comp);
    }
  }
}

/* ------------------------------------------------- assembly + compress --- */

typedef struct { int32_t *ia, *ja; double *val; int64_t nnz, cap; } coo_t;

static void coo_append(coo_t *m, int32_t i, int32_t j, double v) {
  if (m->nnz == m->cap) {
    /* storage grows x1.5 (base_sparse_matrix.f90:3491-3496) */
    int64_t nc = m->cap < 16 ? 64 : (int64_t)(1.5 * (double)m->cap);
    m->ia = (int32_t *)realloc(m->ia, sizeof(int32_t) * (size_t)nc);
    m->ja = (int32_t *)realloc(m->ja, sizeof(int32_t) * (size_t)nc);
    m->val = (double *)realloc(m->val, sizeof(double) * (size_t)nc);
    if (!m->ia || !m->ja || !m->val) { fprintf(stderr, "oracle: COO out of memory\n"); abort(); }
    m->cap = nc;
  }
  m->ia[m->nnz] = i; m->ja[m->nnz] = j; m->val[m->nnz] = v; ++m->nnz;
}

/* sort_and_compress_numeric by rows (base_sparse_matrix.f90:4482-4827):
   stable counting sort on rows, stable column sort inside each row
   (mergesort_link_list, sparse_matrix_utils.f90:136-250), duplicates summed in
   insertion order; then move_from_coo builds irp (csr_sparse_matrix.f90:416-479).
   Rows 1..n; the COO arrays are consumed.  Returns nnz. */
static int64_t coo_compress(coo_t *m, int n, int32_t **irp_out, int32_t **ja_out, double **val_out) {
  int64_t nnz = m->nnz;
  int64_t *cnt = (int64_t *)xcalloc((size_t)n + 2, sizeof(int64_t));
  for (int64_t k = 0; k < nnz; ++k) cnt[m->ia[k]]++;
  int64_t acc = 0;
  for (int r = 1; r <= n; ++r) { int64_t t = cnt[r]; cnt[r] = acc; acc += t; }
  int32_t *ja2 = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
  double *v2 = (double *)xmalloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
  int64_t *rs = (int64_t *)xmalloc(sizeof(int64_t) * ((size_t)n + 2));
  for (int r = 1; r <= n; ++r) rs[r] = cnt[r];
  rs[n + 1] = acc;
  for (int64_t k = 0; k < nnz; ++k) {
    int64_t p = cnt[m->ia[k]]++;
    ja2[p] = m->ja[k]; v2[p] = m->val[k];
  }
  free(m->ia); free(m->ja); free(m->val); free(cnt);
  m->ia = NULL; m->ja = NULL; m->val = NULL; m->nnz = m->cap = 0;

  int32_t *irp = (int32_t *)xmalloc(sizeof(int32_t) * ((size_t)n + 1));
  int32_t *jo = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
  double *vo = (double *)xmalloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
  /* per-row stable merge sort buffers */
  int32_t *tj = NULL; double *tv = NULL; int64_t tcap = 0;
  int64_t k = 0;
  irp[0] = 1;
  for (int r = 1; r <= n; ++r) {
    int64_t b = rs[r], e = (r < n) ? rs[r + 1] : acc, len = e - b;
    if (len > tcap) {
      tcap = len;
      tj = (int32_t *)realloc(tj, sizeof(int32_t) * (size_t)tcap);
      tv = (double *)realloc(tv, sizeof(double) * (size_t)tcap);
    }
    /* bottom-up stable merge sort of (ja2, v2)[b:e) by column */
    int32_t *aj = ja2 + b; double *av = v2 + b;
    for (int64_t w = 1; w < len; w *= 2) {
      for (int64_t lo = 0; lo < len; lo += 2 * w) {
        int64_t mid = lo + w < len ? lo + w : len, hi = lo + 2 * w < len ? lo + 2 * w : len;
        int64_t i = lo, j = mid, o = lo;
        while (i < mid && j < hi) {
          if (aj[j] < aj[i]) { tj[o] = aj[j]; tv[o] = av[j]; ++j; }
          else { tj[o] = aj[i]; tv[o] = av[i]; ++i; }
          ++o;
        }
        while (i < mid) { tj[o] = aj[i]; tv[o] = av[i]; ++i; ++o; }
        while (j < hi) { tj[o] = aj[j]; tv[o] = av[j]; ++j; ++o; }
      }
      memcpy(aj, tj, sizeof(int32_t) * (size_t)len);
      memcpy(av, tv, sizeof(double) * (size_t)len);
    }
    for (int64_t i = 0; i < len; ++i) {
      if (i > 0 && aj[i] == aj[i - 1]) vo[k - 1] = vo[k - 1] + av[i];
      else { jo[k] = aj[i]; vo[k] = av[i]; ++k; }
    }
    irp[r] = (int32_t)(k + 1);
  }
  free(tj); free(tv); free(ja2); free(v2); free(rs);
  *irp_out = irp; *ja_out = jo; *val_out = vo;
  return k;
}

int64_t orc_assemble(const orc_space_t *s, const orc_physics_t *ph, int sym,
                     const double *fixed, int32_t **irp_out, int32_t **ja_out, double **val_out,
                     double **rhs_out) {
  return orc_assemble_cells(s, ph, sym, fixed, 0, s->ncells, irp_out, ja_out, val_out, rhs_out);
}

int64_t orc_assemble_cells(const orc_space_t *s, const orc_physics_t *ph, int sym,
                           const double *fixed, int64_t cell_begin, int64_t cell_end, int32_t **irp_out,
                           int32_t **ja_out, double **val_out, double **rhs_out) {
  int nd = s->nd, n = s->nfree;
  double *elmat = (double *)xmalloc(sizeof(double) * (size_t)(nd * nd));
  double *elvec = (double *)xmalloc(sizeof(double) * (size_t)nd);
  double *rhs = (double *)xcalloc((size_t)n, sizeof(double));
  coo_t m = {0};
  int64_t per = 0;
  for (int i = 0; i < nd; ++i) for (int j = 0; j < nd; ++j) if (!sym || i <= j) ++per;
  if (cell_begin < 0) cell_begin = 0;
  if (cell_end > s->ncells) cell_end = s->ncells;
  m.cap = (cell_end - cell_begin) * (sym ? per : (int64_t)nd * nd);
  if (m.cap < 64) m.cap = 64;
  m.ia = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)m.cap);
  m.ja = (int32_t *)xmalloc(sizeof(int32_t) * (size_t)m.cap);
  m.val = (double *)xmalloc(sizeof(double) * (size_t)m.cap);

  for (int64_t c = cell_begin; c < cell_end; ++c) {
    orc_element(s, ph, c, elmat, elvec);
    const int32_t *g = s->dofs + c * nd;
    /* impose_strong_dirichlet_bcs_conforming (sbm_fe_cell_iterator.i90:1610-1637) */
    if (fixed) {
      for (int j = 0; j < nd; ++j)
        if (g[j] < 0) {
          double gv = fixed[-g[j] - 1];
          for (int i = 0; i < nd; ++i) elvec[i] = elvec[i] - elmat[i + j * nd] * gv;
        }
    }
    /* coo_sparse_matrix_append_dense_values_body (base_sparse_matrix.f90:3469-3513) */
    for (int j = 0; j < nd; ++j) {
      int32_t jas = g[j];
      if (jas < 1 || jas > n) continue;
      for (int i = 0; i < nd; ++i) {
        int32_t ias = g[i];
        if (ias < 1 || ias > n || (sym && ias > jas)) continue;
        coo_append(&m, ias, jas, elmat[i + j * nd]);
      }
    }
    /* serial_scalar_array_add_multiple_entries (serial_scalar_array.f90:247-262) */
    for (int i = 0; i < nd; ++i)
      if (g[i] > 0) rhs[g[i] - 1] = rhs[g[i] - 1] + elvec[i];
  }
  free(elmat);
  free(elvec);

  int64_t k = coo_compress(&m, n, irp_out, ja_out, val_out);
  *rhs_out = rhs;
  return k;
}

/* --------------------------------------------------------- solver --------- */

void orc_spmv(int n, const int32_t *irp, const int32_t *ja, const double *val, int sym,
              const double *x, double *y) {
  if (!sym) {
    for (int r = 0; r < n; ++r) {
      double acc = 0.0;   /* beta = 0 */
      for (int32_t k = irp[r] - 1; k < irp[r + 1] - 1; ++k) acc = acc + 1.0 * x[ja[k] - 1] * val[k];
      y[r] = acc;
    }
  } else {
    for (int r = 0; r < n; ++r) y[r] = 0.0;
    for (int r = 0; r < n; ++r) {
      if (irp[r + 1] - irp[r] <= 0) continue;
      y[r] = y[r] + x[ja[irp[r] - 1] - 1] * val[irp[r] - 1];
      for (int32_t k = irp[r]; k < irp[r + 1] - 1; ++k) {
        int c = ja[k] - 1;
        y[r] = y[r] + 1.0 * x[c] * val[k];
        y[c] = y[c] + 1.0 * x[r] * val[k];
      }
    }
  }
}

void orc_extract_diagonal(int n, const int32_t *irp, const int32_t *ja, const double *val, double *d) {
  for (int r = 0; r < n; ++r) {
    /* binary_search (sparse_matrix_utils.f90:105-134) */
    int lo = irp[r] - 1, hi = irp[r + 1] - 2;
    d[r] = 0.0;
    while (lo <= hi) {
      int mid = (lo + hi) / 2;
      if (ja[mid] == r + 1) { d[r] = val[mid]; break; }
      if (ja[mid] < r + 1) lo = mid + 1; else hi = mid - 1;
    }
  }
}

static double dot(int n, const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
static double nrm2(int n, const double *a) { return sqrt(dot(n, a, a)); }

void orc_cg(int n, const int32_t *irp, const int32_t *ja, const double *val, int sym, int precond,
            int crit, double rtol, double atol, int maxit, const double *b, double *x,
            orc_cg_result_t *res, double *history) {
  double *r = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *z = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *p = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *Ap = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *inv = NULL;
  if (precond == 1) {
    /* jacobi numerical_setup: extract_diagonal + entrywise_invert
       (jacobi_preconditioner.f90:356-389, serial_scalar_array.f90:581-588) */
    inv = (double *)xmalloc(sizeof(double) * (size_t)n);
    orc_extract_diagonal(n, irp, ja, val, inv);
    for (int i = 0; i < n; ++i) inv[i] = 1.0 / inv[i];
  }
#define APPLY_M(src, dst) do { if (inv) { for (int i_ = 0; i_ < n; ++i_) (dst)[i_] = inv[i_] * (src)[i_]; } \
                               else memcpy((dst), (src), sizeof(double) * (size_t)n); } while (0)
  /* cg_solve_body (cg.f90:101-238) */
  double b_nrm_M = 0.0;
  if (crit == ORC_RES_NRMGIVEN_RHS_NRMGIVEN) {
    APPLY_M(b, r);
    b_nrm_M = sqrt(dot(n, b, r));
  }
  orc_spmv(n, irp, ja, val, sym, x, r);
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];          /* r%axpby(1.0, b, -1.0) */
  APPLY_M(r, z);
  double r_z = dot(n, r, z);
  double r_nrm_M = sqrt(r_z);
  memcpy(p, z, sizeof(double) * (size_t)n);
  /* init_convergence_data (cg.f90:279-321) */
  double rhs_test = 0.0, rhs_extra = 0.0;
  switch (crit) {
    case ORC_DELTA_RHS: case ORC_RES_RHS: case ORC_DELTA_RHS_AND_RES_RHS: case ORC_DELTA_DELTA_AND_RES_RHS:
      rhs_test = rtol * nrm2(n, b) + atol;
      if (crit == ORC_DELTA_RHS_AND_RES_RHS || crit == ORC_DELTA_DELTA_AND_RES_RHS) rhs_extra = rhs_test;
      break;
    case ORC_RES_RES: case ORC_DELTA_DELTA_AND_RES_RES:
      rhs_test = rtol * nrm2(n, r) + atol;
      if (crit == ORC_DELTA_DELTA_AND_RES_RES) rhs_extra = rhs_test;
      break;
    case ORC_DELTA_RHS_AND_RES_RES:
      rhs_test = rtol * nrm2(n, b) + atol;
      rhs_extra = rtol * nrm2(n, r) + atol;
      break;
    case ORC_RES_NRMGIVEN_RHS_NRMGIVEN: rhs_test = rtol * b_nrm_M + atol; break;
    case ORC_RES_NRMGIVEN_RES_NRMGIVEN: rhs_test = rtol * r_nrm_M + atol; break;
    default: break;
  }
  int it = 0, conv = 0;
  double est = 0.0;
  int is_delta = (crit == ORC_DELTA_RHS || crit == ORC_DELTA_DELTA || crit == ORC_DELTA_RHS_AND_RES_RES ||
                  crit == ORC_DELTA_RHS_AND_RES_RHS || crit == ORC_DELTA_DELTA_AND_RES_RES ||
                  crit == ORC_DELTA_DELTA_AND_RES_RHS);
  int is_dd = (crit == ORC_DELTA_DELTA || crit == ORC_DELTA_DELTA_AND_RES_RES || crit == ORC_DELTA_DELTA_AND_RES_RHS);
  int has_extra = (crit >= ORC_DELTA_RHS_AND_RES_RES);
  for (;;) {
    ++it;
    orc_spmv(n, irp, ja, val, sym, p, Ap);
    double Ap_p = dot(n, Ap, p);
    double alpha = (Ap_p != 0.0) ? r_z / Ap_p : 0.0;
    for (int i = 0; i < n; ++i) x[i] = x[i] + alpha * p[i];
    for (int i = 0; i < n; ++i) r[i] = r[i] + (-alpha) * Ap[i];
    /* update_convergence_data_and_evaluate_stopping_criteria (cg.f90:323-407) */
    if (it == 1 && is_delta) {
      est = alpha * nrm2(n, p);
      if (is_dd) rhs_test = rtol * est + atol;
    }
    if (crit == ORC_RES_RES || crit == ORC_RES_RHS) est = nrm2(n, r);
    else if (is_delta) { if (it != 1) est = alpha * nrm2(n, p); }
    else est = r_nrm_M;
    if (history) history[it - 1] = est;
    conv = (est <= rhs_test);
    if (has_extra && conv) conv = (nrm2(n, r) <= rhs_extra);
    if (conv || it >= maxit) break;
    APPLY_M(r, z);
    double beta = 1.0 / r_z;
    r_z = dot(n, r, z);
    beta = beta * r_z;
    r_nrm_M = sqrt(r_z);
    /* p%axpby(1.0, z, beta): scal then daxpy (serial_scalar_array.f90:386-394) */
    for (int i = 0; i < n; ++i) { p[i] = beta * p[i]; p[i] = p[i] + 1.0 * z[i]; }
  }
#undef APPLY_M
  res->iterations = it;
  res->converged = conv;
  res->error_estimate = est;
  res->rhs_test = rhs_test;
  free(r); free(z); free(p); free(Ap); free(inv);
}

/* ---- error norms (Integration/ErrorNorms/sbm_error_norms_scalar.i90:67-241) ---- */

static void exact_value(int exact, int d, const double x[3], double *u, double g[3]) {
  g[0] = g[1] = g[2] = 0.0;
  if (exact == ORC_EXACT_LINEAR) {   /* poisson_analytical_functions.f90:77-104 */
    *u = 0.0;
    for (int k = 0; k < d; ++k) {
      *u += x[k];
      g[k] = 1.0;
    }
    return;
  }
  double s[3] = {1.0, 1.0, 1.0}, c[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < d; ++k) {
    s[k] = sin(PI * x[k]);
    c[k] = cos(PI * x[k]);
  }
  *u = s[0] * s[1] * s[2];
  for (int k = 0; k < d; ++k) {
    double v = PI * c[k];
    for (int e = 0; e < d; ++e)
      if (e != k) v *= s[e];
    g[k] = v;
  }
}

void orc_error_norms(const orc_space_t *s, const double *x_free, const double *fixed_values, int exact,
                     double out[4]) {
  const int nd = s->nd, d = s->d;
  double values_norm = 0.0, gradients_norm = 0.0, linf = 0.0;
  double xq[MAXNQ][3], nodal[MAXND];
  cell_geo_t geo;
  for (int64_t c = 0; c < s->ncells; ++c) {   /* fe cell iterator, cells in order (:94-110) */
    qpoint_coords(s, c, xq);
    const cell_geo_t *g = cell_geometry(s, c, &geo);
    const double *detj = g ? g->detj : s->detj;
    const double (*grad)[3][MAXNQ] = g ? g->grad : s->grad;
    const int32_t *cd = s->dofs + c * nd;
    for (int i = 0; i < nd; ++i)   /* fe_cell_function%update: nodal values of the cell */
      nodal[i] = cd[i] > 0 ? x_free[cd[i] - 1] : (fixed_values ? fixed_values[-cd[i] - 1] : 0.0);
    double vc = 0.0, gc = 0.0;
    for (int q = 0; q < s->nq; ++q) {
      double uh = 0.0, guh[3] = {0.0, 0.0, 0.0}, u, gu[3];
      for (int i = 0; i < nd; ++i) {   /* values and gradients at the quadrature point */
        uh += nodal[i] * s->shp[i][q];
        for (int k = 0; k < 3; ++k) guh[k] += nodal[i] * grad[i][k][q];
      }
      exact_value(exact, d, xq[q], &u, gu);
      const double e = u - uh;
      double ge2 = 0.0;
      for (int k = 0; k < 3; ++k) ge2 += (gu[k] - guh[k]) * (gu[k] - guh[k]);
      const double dv = detj[q] * s->qw[q];
      vc += (e * e) * dv;      /* l2_norm, h1_norm (:200-205) */
      gc += ge2 * dv;          /* h1_norm, h1_seminorm (:220-225) */
      if (fabs(e) > linf) linf = fabs(e);
    }
    values_norm += vc;          /* update_norm (error_norms.f90:160-172): sum over cells */
    gradients_norm += gc;
  }
  out[0] = sqrt(values_norm);   /* finalize_norm (error_norms.f90:174-221): square roots */
  out[1] = sqrt(gradients_norm);
  out[2] = sqrt(values_norm + gradients_norm);
  out[3] = linf;
}

/* ------------------------------------------------- multi-field spaces --- */

struct orc_mf {
  int nf, nb, ndt;
  orc_space_t *f[ORC_MF_MAXF];
  int block[ORC_MF_MAXF];      /* 1-based block of each field (block_layout_t%field_id_to_block_id) */
  int eoff[ORC_MF_MAXF + 1];   /* first row of each field in the cell's element matrix */
  int ndofs_block[ORC_MF_MAXF + 1];
  int32_t *dofs[ORC_MF_MAXF];  /* cells x nd_f, block-local free ids and global fixed ids */
};

/* serial_fe_space_t with several fields on one triangulation
   (sbm_serial_fe_space.i90:40-140 create, block_layout_t%create
   Integration/FEOperator/block_layout.f90:66-117).  Strong Dirichlet ids: the
   sweep runs field by field over all cells with one counter
   (sbm_serial_fe_space.i90:584-640), so field f's fixed ids follow those of
   fields 1..f-1.  Free ids: each field has its own counter, started at the
   number of free DOFs of the earlier fields of its block
   (sbm_serial_fe_space.i90:2676-2682), and numbers first touch by cells as
   one field alone (:2684-2700).  Every field is integrated with the
   quadrature of the highest order in the cell (:1891-2006). */
orc_mf_t *orc_mf_create(int nfields, const orc_problem_t *fields, const int *field_blocks) {
  if (nfields < 1 || nfields > ORC_MF_MAXF) return NULL;
  int maxo = 0, ndt = 0;
  for (int i = 0; i < nfields; ++i) {
    const orc_problem_t *a = &fields[i], *b = &fields[0];
    if (a->dim != b->dim || a->ncomp != 1 || field_blocks[i] < 1 || field_blocks[i] > nfields) return NULL;
    for (int k = 0; k < 3; ++k) if (k < a->dim && a->n[k] != b->n[k]) return NULL;
    for (int k = 0; k < 2 * a->dim; ++k) if (a->lim[k] != b->lim[k]) return NULL;
    if (a->order > maxo) maxo = a->order;
  }
  orc_mf_t *m = (orc_mf_t *)xcalloc(1, sizeof(orc_mf_t));
  m->nf = nfields;
  for (int i = 0; i < nfields; ++i) {
    m->f[i] = space_create(&fields[i], maxo + 1);
    if (!m->f[i]) { orc_mf_free(m); return NULL; }
    m->block[i] = field_blocks[i];
    if (field_blocks[i] > m->nb) m->nb = field_blocks[i];
    m->eoff[i] = ndt;
    ndt += m->f[i]->nd;
  }
  m->eoff[nfields] = ndt;
  m->ndt = ndt;
  int free_off[ORC_MF_MAXF] = {0}, fixed_off = 0;
  for (int i = 0; i < nfields; ++i) {   /* offsets in field order */
    free_off[i] = m->ndofs_block[m->block[i]];
    m->ndofs_block[m->block[i]] += m->f[i]->nfree;
  }
  for (int i = 0; i < nfields; ++i) {
    const orc_space_t *s = m->f[i];
    size_t cnt = (size_t)(s->ncells * s->nd);
    m->dofs[i] = (int32_t *)xmalloc(sizeof(int32_t) * (cnt ? cnt : 1));
    for (size_t k = 0; k < cnt; ++k) {
      const int32_t g = s->dofs[k];
      m->dofs[i][k] = g > 0 ? g + free_off[i] : g - fixed_off;
    }
    fixed_off += s->nfixed;
  }
  return m;
}

void orc_mf_free(orc_mf_t *m) {
  if (!m) return;
  for (int i = 0; i < m->nf; ++i) { orc_space_free(m->f[i]); free(m->dofs[i]); }
  free(m);
}

int orc_mf_num_blocks(const orc_mf_t *m) { return m->nb; }

void orc_mf_field_sizes(const orc_mf_t *m, int field, int *nd, int *n_free, int *n_fixed) {
  const orc_space_t *s = m->f[field];
  *nd = s->nd; *n_free = s->nfree; *n_fixed = s->nfixed;
}

int orc_mf_block_size(const orc_mf_t *m, int block) { return m->ndofs_block[block]; }

void orc_mf_get_dofs(const orc_mf_t *m, int field, int32_t *out) {
  const orc_space_t *s = m->f[field];
  memcpy(out, m->dofs[field], sizeof(int32_t) * (size_t)(s->ncells * s->nd));
}

/* Element of the coupled diffusion-reaction plugin on field-major element
   rows (field 1's DOFs, then field 2's, ...), the q loop outermost and then
   i and j as the Poisson plugin (poisson_discrete_integration.f90:116-129).
   The bilinear form sum_f kappa_f (grad u_f, grad v_f) + sigma sum_{f<g}
   (u_f - u_g, v_f - v_g) (symmetric positive definite with Dirichlet data on
   some field) gives
     (f, f): kappa_f grad phi_i . grad phi_j + (nf - 1) sigma phi_i phi_j
     (f, g): -sigma phi_i psi_j           (f != g)
     elvec_f: f_f phi_i
   elmat column-major (ndt x ndt). */
void orc_mf_element(const orc_mf_t *m, const orc_mf_physics_t *ph, int64_t cell, double *elmat,
                    double *elvec) {
  (void)cell;   /* constant coefficients on a uniform mesh */
  const int ndt = m->ndt;
  const orc_space_t *s0 = m->f[0];
  for (int i = 0; i < ndt * ndt; ++i) elmat[i] = 0.0;
  for (int i = 0; i < ndt; ++i) elvec[i] = 0.0;
  for (int q = 0; q < s0->nq; ++q) {
    const double dv = s0->detj[q] * s0->qw[q];
    for (int fa = 0; fa < m->nf; ++fa) {
      const orc_space_t *A = m->f[fa];
      for (int i = 0; i < A->nd; ++i) {
        for (int fb = 0; fb < m->nf; ++fb) {
          const orc_space_t *B = m->f[fb];
          for (int j = 0; j < B->nd; ++j) {
            double v;
            if (fa == fb) {
              double dot = 0.0;
              for (int k = 0; k < 3; ++k) dot = dot + A->grad[j][k][q] * A->grad[i][k][q];
              v = ph->kappa[fa] * dot + (m->nf - 1) * ph->sigma * (A->shp[i][q] * A->shp[j][q]);
            } else {
              v = -ph->sigma * (A->shp[i][q] * B->shp[j][q]);
            }
            elmat[(m->eoff[fa] + i) + (size_t)(m->eoff[fb] + j) * ndt] += dv * v;
          }
        }
        elvec[m->eoff[fa] + i] += dv * ph->f[fa] * A->shp[i][q];
      }
    }
  }
}

/* Block (ib, jb) of the block_sparse_matrix_t (1-based blocks): the cells in
   order, and per cell element_block_sparse_matrix_assembly
   (Integration/FEOperator/block_sparse_assembler.f90:182-222): row fields in
   order, column fields in order, each coupled pair's dense sub-block
   inserted (j outer, i inner; columns or rows outside the block's range and,
   with symmetric storage of a diagonal block, row > col skipped:
   base_sparse_matrix.f90:3469-3513), then sort and compress.  field_coupling
   is nf x nf, [f * nf + g] nonzero when field f's rows take field g's columns
   (block_layout_t field_coupling). Homogeneous Dirichlet: no lifts. */
int64_t orc_mf_assemble_block(const orc_mf_t *m, const orc_mf_physics_t *ph, const int *field_coupling, int ib,
                              int jb, int symmetric_storage, int32_t **irp, int32_t **ja, double **val) {
  const int ndt = m->ndt, nr = m->ndofs_block[ib], nc = m->ndofs_block[jb];
  const int sym = symmetric_storage && ib == jb;
  double *elmat = (double *)xmalloc(sizeof(double) * (size_t)(ndt * ndt));
  double *elvec = (double *)xmalloc(sizeof(double) * (size_t)ndt);
  coo_t c = {0};
  const int64_t ncells = m->f[0]->ncells;
  for (int64_t cell = 0; cell < ncells; ++cell) {
    orc_mf_element(m, ph, cell, elmat, elvec);
    for (int fa = 0; fa < m->nf; ++fa) {
      if (m->block[fa] != ib) continue;
      const int nda = m->f[fa]->nd;
      const int32_t *ga = m->dofs[fa] + cell * nda;
      for (int fb = 0; fb < m->nf; ++fb) {
        if (m->block[fb] != jb || !field_coupling[fa * m->nf + fb]) continue;
        const int ndb = m->f[fb]->nd;
        const int32_t *gb = m->dofs[fb] + cell * ndb;
        for (int j = 0; j < ndb; ++j) {
          const int32_t jas = gb[j];
          if (jas < 1 || jas > nc) continue;
          for (int i = 0; i < nda; ++i) {
            const int32_t ias = ga[i];
            if (ias < 1 || ias > nr || (sym && ias > jas)) continue;
            coo_append(&c, ias, jas, elmat[(m->eoff[fa] + i) + (size_t)(m->eoff[fb] + j) * ndt]);
          }
        }
      }
    }
  }
  free(elmat);
  free(elvec);
  return coo_compress(&c, nr, irp, ja, val);
}

/* Block ib of the serial_block_array_t right-hand side
   (element_serial_block_array_assembly, block_sparse_assembler.f90:157-180):
   cells in order, fields of the block in order. */
void orc_mf_assemble_rhs(const orc_mf_t *m, const orc_mf_physics_t *ph, int ib, double *rhs) {
  const int ndt = m->ndt, nr = m->ndofs_block[ib];
  double *elmat = (double *)xmalloc(sizeof(double) * (size_t)(ndt * ndt));
  double *elvec = (double *)xmalloc(sizeof(double) * (size_t)ndt);
  for (int i = 0; i < nr; ++i) rhs[i] = 0.0;
  const int64_t ncells = m->f[0]->ncells;
  for (int64_t cell = 0; cell < ncells; ++cell) {
    orc_mf_element(m, ph, cell, elmat, elvec);
    for (int fa = 0; fa < m->nf; ++fa) {
      if (m->block[fa] != ib) continue;
      const int nda = m->f[fa]->nd;
      const int32_t *ga = m->dofs[fa] + cell * nda;
      for (int i = 0; i < nda; ++i)
        if (ga[i] > 0) rhs[ga[i] - 1] = rhs[ga[i] - 1] + elvec[m->eoff[fa] + i];
    }
  }
  free(elmat);
  free(elvec);
}

/* ------------------------------------------ all-cores timing baseline --- */

/* Jacobi-PCG with the RES_RES criterion on a full-storage CSR matrix, its
   loops spread over the host's threads (OpenMP).  Not a restatement of the
   reference (whose serial path has no threading and applies symmetric
   storage): the same recurrence as orc_cg, used only to time an all-cores CPU
   figure beside the one-core oracle. */
void orc_cg_threads(int n, const int32_t *irp, const int32_t *ja, const double *val, double rtol, int maxit,
                    const double *b, double *x, orc_cg_result_t *res) {
  double *r = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *z = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *p = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *q = (double *)xmalloc(sizeof(double) * (size_t)n);
  double *inv = (double *)xmalloc(sizeof(double) * (size_t)n);
  double rz = 0.0, rr = 0.0;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double d = 0.0, acc = 0.0;
    for (int32_t k = irp[i] - 1; k < irp[i + 1] - 1; ++k) {
      if (ja[k] - 1 == i) d = val[k];
      acc += val[k] * x[ja[k] - 1];
    }
    inv[i] = 1.0 / d;
    r[i] = b[i] - acc;
  }
#pragma omp parallel for schedule(static) reduction(+ : rz, rr)
  for (int i = 0; i < n; ++i) {
    z[i] = inv[i] * r[i];
    p[i] = z[i];
    rz += r[i] * z[i];
    rr += r[i] * r[i];
  }
  const double thr = rtol * sqrt(rr);
  int it = 0, conv = 0;
  double est = sqrt(rr);
  while (it < maxit) {
    ++it;
    double pq = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : pq)
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int32_t k = irp[i] - 1; k < irp[i + 1] - 1; ++k) acc += val[k] * p[ja[k] - 1];
      q[i] = acc;
      pq += acc * p[i];
    }
    const double alpha = pq != 0.0 ? rz / pq : 0.0;
    double rr2 = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr2)
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      rr2 += r[i] * r[i];
    }
    est = sqrt(rr2);
    if (est <= thr) { conv = 1; break; }
    double rz2 = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rz2)
    for (int i = 0; i < n; ++i) {
      z[i] = inv[i] * r[i];
      rz2 += r[i] * z[i];
    }
    const double beta = rz2 / rz;
    rz = rz2;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  res->iterations = it;
  res->converged = conv;
  res->error_estimate = est;
  res->rhs_test = thr;
  free(r); free(z); free(p); free(q); free(inv);
}
