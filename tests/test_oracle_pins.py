"""Pins of the CPU oracle against the reference's known answers and against
what the mathematics fixes.  CPU only.

Each check is chosen so that a plausible slip (dropped term, wrong sign or
index, transposed operand, wrong numbering order) fails it.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- quadrature ---

def test_gauss_legendre_matches_reference_tables():
    """test_quadratures.f90 tables, 1..20 points, tolerance 1e-10 (test_quadratures.f90:505-519)."""
    d = json.load(open(os.path.join(GOLD, "gauss_legendre_tables.json")))
    for n_str, t in d["tables"].items():
        n = int(n_str)
        x, w = O.gauss_legendre(n)
        ref_x = np.sort(np.array(t["x"]))
        order = np.argsort(np.array(t["x"]))
        ref_w = np.array(t["w"])[order]
        assert np.all(np.diff(x) > 0), "abscissae must be ascending"
        assert np.max(np.abs(x - ref_x)) < d["tolerance"], n
        assert np.max(np.abs(w - ref_w)) < d["tolerance"], n


def test_gauss_legendre_closed_forms_and_exactness():
    x, w = O.gauss_legendre(2)
    assert np.allclose(x, [-1 / np.sqrt(3), 1 / np.sqrt(3)], rtol=0, atol=1e-15)
    assert np.allclose(w, [1, 1], rtol=0, atol=1e-15)
    x, w = O.gauss_legendre(3)
    assert np.allclose(x, [-np.sqrt(0.6), 0, np.sqrt(0.6)], rtol=0, atol=1e-15)
    assert np.allclose(w, [5 / 9, 8 / 9, 5 / 9], rtol=0, atol=1e-15)
    for n in range(1, 12):   # exact for polynomials of degree 2n-1
        x, w = O.gauss_legendre(n)
        for deg in range(2 * n):
            exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
            assert abs(np.dot(w, x ** deg) - exact) < 1e-13


# ------------------------------------------------------------ shape functions ---

@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_shape_function_invariants(dim, order):
    sp = O.Space(O.make_problem(dim, [2] * dim, order))
    qp, qw, shp, dshp = sp.tables()
    assert np.allclose(shp.sum(axis=1), 1.0, atol=1e-14)            # partition of unity
    assert np.allclose(dshp.sum(axis=1), 0.0, atol=1e-13)           # gradients sum to zero
    assert abs(qw.sum() - 2.0 ** dim) < 1e-13
    # tensor structure: node a = (a_x, a_y, a_z) x fastest; value = prod of 1-D Lagrange
    t = np.linspace(-1, 1, order + 1)
    for q in range(len(qw)):
        for a in range(shp.shape[1]):
            idx = np.unravel_index(a, [order + 1] * dim, order="F")
            v = 1.0
            for k in range(dim):
                others = [t[m] for m in range(order + 1) if m != idx[k]]
                v *= np.prod([(qp[q, k] - o) / (t[idx[k]] - o) for o in others])
            assert abs(shp[q, a] - v) < 1e-14
    # quadrature points: first index fastest
    assert qp[1, 0] > qp[0, 0] and (dim < 2 or qp[1, 1] == qp[0, 1])


# ------------------------------------------------------------- numbering -------

def test_numbering_2d_q1_3x3_hand_derived():
    """First touch, Dirichlet first (negative), x-fastest cells and vertices."""
    sp = O.Space(O.make_problem(2, (3, 3), 1))
    expected = np.array([[-1, -2, -3, 1], [-2, -4, 1, 2], [-4, -5, 2, -6],
                         [-3, 1, -7, 3], [1, 2, 3, 4], [2, -6, 4, -8],
                         [-7, 3, -9, -10], [3, 4, -10, -11], [4, -8, -11, -12]])
    assert np.array_equal(sp.dofs(), expected)
    assert (sp.n_free, sp.n_fixed) == (4, 12)


def test_numbering_2d_q2_2x2_hand_derived():
    """Cell interior first, then vertices, then edges (x-edges y=0,1; y-edges x=0,1)."""
    sp = O.Space(O.make_problem(2, (2, 2), 2))
    d = sp.dofs()
    # cell 0 nodes (x fastest over 3x3): interior (1,1)=1, vertex (2,2)=2, x-edge y=1 -> 3, y-edge x=1 -> 4
    assert d[0, 4] == 1 and d[0, 8] == 2 and d[0, 7] == 3 and d[0, 5] == 4
    assert d[1, 4] == 5 and d[1, 7] == 6 and d[1, 6] == 2 and d[1, 3] == 4
    assert d[2, 4] == 7 and d[2, 5] == 8 and d[2, 2] == 2 and d[2, 1] == 3
    assert d[3, 4] == 9 and d[3, 0] == 2 and d[3, 1] == 6 and d[3, 3] == 8
    # fixed ids of cell 0: vertices 0,1,2 -> -1,-2,-3 ; x-edge y=0 -> -4 ; y-edge x=0 -> -5
    assert [d[0, 0], d[0, 2], d[0, 6], d[0, 1], d[0, 3]] == [-1, -2, -3, -4, -5]
    assert (sp.n_free, sp.n_fixed) == (9, 16)


def test_numbering_3d_q1_2x2x2_and_interior_lexicographic():
    sp = O.Space(O.make_problem(3, (2, 2, 2), 1))
    d = sp.dofs()
    assert sp.n_free == 1 and d[0, 7] == 1 and d[7, 0] == 1
    # all-Dirichlet Q1: free ids are lexicographic over interior vertices
    n = 5
    sp = O.Space(O.make_problem(3, (n, n, n), 1))
    d = sp.dofs()
    for c in range(sp.num_cells):
        i, j, k = c % n, (c // n) % n, c // (n * n)
        if i < n - 1 and j < n - 1 and k < n - 1:
            assert d[c, 7] == 1 + i + (n - 1) * j + (n - 1) ** 2 * k


def test_numbering_vector_q1_interleaved_and_clamped_face():
    """Vector Q1: components of a vertex are consecutive ids; clamped x=0 face only."""
    sp = O.Space(O.make_problem(3, (2, 1, 1), 1, ncomp=3, bc=O.BC_X0))
    d = sp.dofs()
    assert sp.n_fixed == 3 * 4 and sp.n_free == 3 * 8
    # cell 0: vertices with x=0 (local 0,2,4,6) fixed; vertex 1 first free -> ids 1,2,3 (comp-major local index)
    nodes = 8
    assert [d[0, c * nodes + 1] for c in range(3)] == [1, 2, 3]
    assert [d[0, c * nodes + 3] for c in range(3)] == [4, 5, 6]
    assert all(d[0, c * nodes + a] < 0 for c in range(3) for a in (0, 2, 4, 6))
    assert [d[0, c * nodes + 0] for c in range(3)] == [-1, -2, -3]


# ----------------------------------------------------------------- sizes -------

@pytest.mark.parametrize("dim,n,order", [(2, 64, 1), (3, 16, 1), (2, 8, 2), (3, 6, 2)])
def test_pattern_sizes_closed_form(dim, n, order):
    sp = O.Space(O.make_problem(dim, [n] * dim, order))
    irp, ja, val, rhs = sp.assemble(O.make_physics(), symmetric_storage=False)
    m = order * n - 1                         # interior nodes per direction
    s = 3 * m - 2 if order == 1 else None
    if order == 1:
        assert sp.n_free == m ** dim
        assert len(ja) == s ** dim
    else:
        assert sp.n_free == m ** dim
    irp_u, ja_u, _, _ = sp.assemble(O.make_physics(), symmetric_storage=True)
    assert len(ja_u) == (len(ja) + sp.n_free) // 2
    rows = np.repeat(np.arange(1, sp.n_free + 1), np.diff(irp_u))
    assert np.all(ja_u >= rows)
    assert np.all(ja_u[irp_u[:-1] - 1] == np.arange(1, sp.n_free + 1))    # diagonal first
    for r in range(sp.n_free):
        seg = ja[irp[r] - 1: irp[r + 1] - 1]
        assert np.all(np.diff(seg) > 0)


# ----------------------------------------------------------------- values ------

def test_q1_stencils_closed_form():
    h = 1 / 8
    sp = O.Space(O.make_problem(2, (8, 8), 1))
    irp, ja, val, rhs = sp.assemble(O.make_physics(f_kind=O.F_CONST, f0=1.0), symmetric_storage=False)
    r = 3 * 7 + 3
    seg = val[irp[r] - 1: irp[r + 1] - 1]
    assert np.allclose(sorted(seg), [-1 / 3] * 4 + [-1 / 3] * 4 + [8 / 3], atol=1e-14)
    assert abs(rhs[r] - h * h) < 1e-15           # f = 1: rhs = int phi = h^2
    sp = O.Space(O.make_problem(3, (8, 8, 8), 1))
    irp, ja, val, rhs = sp.assemble(O.make_physics(f_kind=O.F_CONST, f0=2.0), symmetric_storage=False)
    r = 3 * 49 + 3 * 7 + 3
    cols = ja[irp[r] - 1: irp[r + 1] - 1] - 1
    seg = val[irp[r] - 1: irp[r + 1] - 1]
    for c, v in zip(cols, seg):
        di = np.array(np.unravel_index(c, (7, 7, 7), order="F")) - np.array(np.unravel_index(r, (7, 7, 7), order="F"))
        nz = np.count_nonzero(di)
        expect = {0: 8 * h / 3, 1: 0.0, 2: -h / 6, 3: -h / 12}[nz]
        assert abs(v - expect) < 1e-15, (di, v, expect)
    assert abs(rhs[r] - 2.0 * h ** 3) < 1e-15


def _kron_ref(dim, order, h):
    if order == 1:
        K1 = np.array([[1, -1], [-1, 1]]) / h
        M1 = np.array([[2, 1], [1, 2]]) * h / 6
    else:
        K1 = np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / (3 * h)
        M1 = np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) * h / 30
    mats = []
    for k in range(dim):
        term = np.ones((1, 1))
        for j in reversed(range(dim)):       # node index x fastest -> kron(z, y, x)
            term = np.kron(term, K1 if j == k else M1)
        mats.append(term)
    return sum(mats)


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_element_matrix_tensor_closed_form(dim, order):
    """Laplace element on a box = sum of Kronecker products of exact 1-D stiffness
    and mass matrices (Gauss with order+1 points integrates them exactly)."""
    n = 4
    h = 1.0 / n
    sp = O.Space(O.make_problem(dim, [n] * dim, order))
    em, ev = sp.element(O.make_physics(f_kind=O.F_CONST, f0=3.0), 5)
    ref = _kron_ref(dim, order, h)
    assert np.max(np.abs(em - ref)) < 1e-13 * np.max(np.abs(ref))
    # f = const: elvec = f * int N_a = f * tensor product of 1-D integrals
    w1 = np.array([0.5, 0.5]) * h if order == 1 else np.array([1, 4, 1]) / 6 * h
    wt = np.ones(1)
    for _ in range(dim):
        wt = np.kron(w1, wt)
    assert np.allclose(ev, 3.0 * wt, rtol=0, atol=1e-15)


def test_coefficient_scales_and_manufactured_source_signs():
    """k const scales the element linearly; the manufactured source for u = sin sin is positive inside."""
    sp = O.Space(O.make_problem(2, (4, 4), 1))
    e1, _ = sp.element(O.make_physics(k0=1.0), 5)
    e3, _ = sp.element(O.make_physics(k0=3.0), 5)
    assert np.allclose(e3, 3 * e1, rtol=1e-15, atol=0)
    _, ev = sp.element(O.make_physics(f_kind=O.F_MANUFACTURED), 5)
    assert np.all(ev > 0)


def test_elasticity_element_energies_and_rigid_modes():
    E, nu = 7.0, 0.3
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    n = 4
    h = 1.0 / n
    sp = O.Space(O.make_problem(3, (n, n, n), 1, ncomp=3, bc=O.BC_X0))
    em, ev = sp.element(O.make_physics(kind=O.PHYS_ELASTICITY, E=E, nu=nu, body=(0, 0, -1)), 0)
    assert np.allclose(em, em.T, atol=1e-14)
    xv = np.array([[(a >> 0) & 1, (a >> 1) & 1, (a >> 2) & 1] for a in range(8)], float) * h

    def field(f):   # nodal values, component-major local index (comp-1)*8 + node
        return np.concatenate([np.array([f(x)[c] for x in xv]) for c in range(3)])

    for f in [lambda x: (1, 0, 0), lambda x: (0, 1, 0), lambda x: (0, 0, 1),
              lambda x: (-x[1], x[0], 0), lambda x: (0, -x[2], x[1]), lambda x: (x[2], 0, -x[0])]:
        u = field(f)
        assert np.max(np.abs(em @ u)) < 1e-13          # rigid body modes carry no energy
    vol = h ** 3
    u = field(lambda x: (x[0], 0, 0))                   # eps_xx = 1
    assert abs(u @ em @ u - (lam + 2 * mu) * vol) < 1e-13
    u = field(lambda x: (x[1], 0, 0))                   # eps_xy = 1/2
    assert abs(u @ em @ u - mu * vol) < 1e-13
    w = np.linalg.eigvalsh(em)
    assert np.sum(np.abs(w) < 1e-12) == 6 and np.all(w > -1e-12)
    # body force (0,0,-1): z-components get -vol/8, others 0
    assert np.allclose(ev[16:], -vol / 8, atol=1e-16) and np.allclose(ev[:16], 0)


# --------------------------------------------------------------- solver -------

def _full_from_upper(irp, ja, val):
    import scipy.sparse as sps
    n = len(irp) - 1
    rows = np.repeat(np.arange(n), np.diff(irp))
    U = sps.csr_matrix((val, (rows, ja - 1)), shape=(n, n))
    return (U + U.T - sps.diags(U.diagonal())).tocsr()


@pytest.mark.parametrize("dim,n,order,prec", [(2, 12, 1, "jacobi"), (3, 6, 1, "identity"), (2, 5, 2, "jacobi"),
                                              (3, 4, 2, "jacobi")])
def test_cg_matches_direct_solve(dim, n, order, prec):
    import scipy.sparse.linalg as spla
    sp = O.Space(O.make_problem(dim, [n] * dim, order))
    ph = O.make_physics(k_kind=O.K_SIN3, f_kind=O.F_MANUFACTURED)
    irp, ja, val, rhs = sp.assemble(ph, symmetric_storage=True)
    A = _full_from_upper(irp, ja, val)
    x_ref = spla.spsolve(A.tocsc(), rhs)
    x, info = O.cg(irp, ja, val, rhs, True, precond=prec, rtol=1e-12, maxit=5000)
    assert info["converged"]
    assert np.linalg.norm(x - x_ref) <= 1e-10 * np.linalg.norm(x_ref)
    # the symmetric-storage SpMV equals the expanded full matrix product
    v = np.random.default_rng(1).standard_normal(len(rhs))
    assert np.allclose(O.spmv(irp, ja, val, v, True), A @ v, rtol=1e-13, atol=1e-14)
    irp_f, ja_f, val_f, rhs_f = sp.assemble(ph, symmetric_storage=False)
    assert np.allclose(O.spmv(irp_f, ja_f, val_f, v, False), A @ v, rtol=1e-13, atol=1e-14)
    assert np.array_equal(rhs, rhs_f)


@pytest.mark.parametrize("dim,order", [(2, 1), (3, 1), (2, 2), (3, 2)])
def test_linear_solution_reproduced_exactly(dim, order):
    """test_poisson: u = x + y (+ z), f = 0, g = u  ->  FE solution equals u at the
    nodes (test_poisson_driver.f90:562-600)."""
    n = 5
    sp = O.Space(O.make_problem(dim, [n] * dim, order))
    ph = O.make_physics(f_kind=O.F_CONST, f0=0.0, g_kind=O.G_LINEAR)
    g = sp.fixed_values(ph)
    irp, ja, val, rhs = sp.assemble(ph, True, fixed_values=g)
    x, info = O.cg(irp, ja, val, rhs, True, precond="identity", rtol=1e-12, maxit=5000)
    d = sp.dofs()
    # nodal coordinates from the cell ijk and local node index
    p1 = order + 1
    for c in range(sp.num_cells):
        ijk = np.unravel_index(c, [n] * dim, order="F")
        for a in range(d.shape[1]):
            if d[c, a] <= 0:
                continue
            m = np.unravel_index(a, [p1] * dim, order="F")
            u = sum((ijk[k] + m[k] / order) / n for k in range(dim))
            assert abs(x[d[c, a] - 1] - u) < 1e-9


def test_cg_reference_semantics():
    """Jacobi-PCG on a diagonal matrix converges in one iteration; maxit=1 stops
    unconverged; DELTA criteria see alpha*||p||; RES_RES tests ||r_i|| <= rtol ||r_0||."""
    n = 6
    irp = np.arange(1, n + 2, dtype=np.int32)
    ja = np.arange(1, n + 1, dtype=np.int32)
    val = np.arange(1, n + 1, dtype=float)
    b = np.ones(n)
    x, info = O.cg(irp, ja, val, b, True, precond="jacobi", rtol=1e-12)
    assert info["iterations"] == 1 and np.allclose(x, 1 / val)
    x, info = O.cg(irp, ja, val, b, True, precond="identity", rtol=1e-14, maxit=1)
    assert info["iterations"] == 1 and not info["converged"]
    x, info = O.cg(irp, ja, val, b, True, precond="identity", rtol=1e-12, maxit=100)
    assert info["iterations"] == n      # n distinct eigenvalues
    sp = O.Space(O.make_problem(2, (10, 10), 1))
    irp, ja, val, rhs = sp.assemble(O.make_physics(f_kind=O.F_CONST, f0=1.0), True)
    for crit in O.CRITERIA:
        x, info = O.cg(irp, ja, val, rhs, True, precond="jacobi", criterion=crit, rtol=1e-8, maxit=500)
        assert info["converged"], crit
        assert info["error_estimate"] <= info["rhs_test"]
    x, info = O.cg(irp, ja, val, rhs, True, precond="jacobi", criterion="RES_RES", rtol=1e-8, history=True)
    r0 = np.linalg.norm(rhs)
    assert abs(info["rhs_test"] - 1e-8 * r0) < 1e-20
    assert info["history"][-1] <= 1e-8 * r0 < info["history"][-2]


@pytest.mark.parametrize("dim,order,n", [(2, 1, 32), (2, 2, 16), (3, 1, 16), (3, 2, 8)])
def test_error_norms_of_zero_match_closed_forms(dim, order, n):
    """u_h = 0: ||u||_L2 = 2^(-d/2) and |u|_H1 = sqrt(d pi^2 / 2^d) for
    u = prod sin(pi x_d); composite Gauss quadrature integrates sin^2 on uniform
    cells exactly, so the norms hold to rounding; linfty is below max u = 1."""
    sp = O.Space(O.make_problem(dim, [n] * dim, order))
    r = sp.error_norms(np.zeros(sp.n_free))
    assert abs(r["l2_norm"] - 0.5 ** (dim / 2)) < 1e-12
    assert abs(r["h1_seminorm"] - np.sqrt(dim * np.pi ** 2 / 2 ** dim)) < 1e-12
    assert abs(r["h1_norm"] - np.hypot(r["l2_norm"], r["h1_seminorm"])) < 1e-14
    assert 0.99 < r["linfty_norm"] < 1.0


@pytest.mark.parametrize("dim,order", [(2, 1), (3, 1), (2, 2)])
def test_error_norms_vanish_for_the_linear_solution(dim, order):
    """test_poisson checks the error norms of u = x + y (+ z) (test_poisson_driver.f90:562-600):
    the FE space holds u, so every norm of u - u_h is at rounding level."""
    n = 4
    sp = O.Space(O.make_problem(dim, [n] * dim, order))
    ph = O.make_physics(f_kind=O.F_CONST, f0=0.0, g_kind=O.G_LINEAR)
    g = sp.fixed_values(ph)
    irp, ja, val, rhs = sp.assemble(ph, True, fixed_values=g)
    x, info = O.cg(irp, ja, val, rhs, True, precond="identity", rtol=1e-14, maxit=5000)
    r = sp.error_norms(x, g, exact="linear")
    assert max(r.values()) < 1e-10
    r0 = sp.error_norms(np.zeros(sp.n_free), g, exact="linear")   # a wrong u_h is seen
    assert r0["l2_norm"] > 0.1 and r0["h1_seminorm"] > 0.1


@pytest.mark.parametrize("order,l2_rate,h1_rate", [(1, 4.0, 2.0), (2, 8.0, 4.0)])
def test_error_norms_converge_at_the_theoretical_rates(order, l2_rate, h1_rate):
    """-lap u = f with u = sin(pi x) sin(pi y): halving h divides the L2 error by
    2^(p+1) and the H1 seminorm error by 2^p."""
    errs = []
    for n in (8, 16, 32):
        sp = O.Space(O.make_problem(2, (n, n), order))
        ph = O.make_physics(k_kind=O.K_CONST, k0=1.0, f_kind=O.F_MANUFACTURED)
        irp, ja, val, rhs = sp.assemble(ph, False)
        x, info = O.cg(irp, ja, val, rhs, False, precond="jacobi", rtol=1e-13, maxit=100000)
        errs.append(sp.error_norms(x))
    for a, b in zip(errs, errs[1:]):
        assert abs(a["l2_norm"] / b["l2_norm"] / l2_rate - 1) < 0.02
        assert abs(a["h1_seminorm"] / b["h1_seminorm"] / h1_rate - 1) < 0.02


# ------------------------------------------- variable-coefficient branches -----
# The 3-D benchmark configurations (128^3 Q1, 64^3 Q2) run the oracle's k(x_q) and manufactured-source
# code (k = 1 + 0.5 sin(2 pi x) sin(2 pi y) sin(2 pi z), f = -div(k grad u)); the
# pins below check those branches without the oracle's own formulas.

def _lagrange_1d(p, x):
    """Values and derivatives of the p+1 Lagrange polynomials through equispaced
    nodes of [-1, 1] at the points x (shape (len(x), p+1))."""
    t = np.linspace(-1.0, 1.0, p + 1)
    v = np.ones((len(x), p + 1))
    dv = np.zeros((len(x), p + 1))
    for m in range(p + 1):
        others = [t[j] for j in range(p + 1) if j != m]
        den = np.prod([t[m] - o for o in others])
        v[:, m] = np.prod([x - o for o in others], axis=0) / den if others else 1.0
        for skip in range(len(others)):
            dv[:, m] += np.prod([x - o for j, o in enumerate(others) if j != skip], axis=0) / den
    return v, dv


def _k_sin3(x):
    return 1.0 + 0.5 * np.prod(np.sin(2 * np.pi * x), axis=-1)


def _source_fd(x, eps=1e-3):
    """f = -div(k grad u) for u = prod sin(pi x_d), with the divergence taken by a
    fourth-order central difference of the flux k du/dx_k (not by the closed form)."""
    d = x.shape[-1]

    def flux(y, k):
        g = np.pi * np.cos(np.pi * y[..., k])
        for j in range(d):
            if j != k:
                g = g * np.sin(np.pi * y[..., j])
        return _k_sin3(y) * g

    f = np.zeros(x.shape[:-1])
    for k in range(d):
        e = np.zeros(d)
        e[k] = eps
        f -= (-flux(x + 2 * e, k) + 8 * flux(x + e, k) - 8 * flux(x - e, k) + flux(x - 2 * e, k)) / (12 * eps)
    return f


def _element_numpy(dim, order, n, lim, cell, k_fun, f_fun):
    """Element matrix and vector of one box cell from numpy's Gauss-Legendre rule
    and an explicit tensor Lagrange basis (nodes x fastest), k and f evaluated at
    the mapped points x = x0 + (xi + 1) h / 2."""
    xg, wg = np.polynomial.legendre.leggauss(order + 1)
    v1, dv1 = _lagrange_1d(order, xg)
    ijk = np.unravel_index(cell, n, order="F")
    h = np.array([(lim[2 * k + 1] - lim[2 * k]) / n[k] for k in range(dim)])
    x0 = np.array([lim[2 * k] + ijk[k] * h[k] for k in range(dim)])
    p1 = order + 1
    nd = p1 ** dim
    em = np.zeros((nd, nd))
    ev = np.zeros(nd)
    for qi in np.ndindex(*([p1] * dim)):
        xi = np.array([xg[qi[k]] for k in range(dim)])
        w = np.prod([wg[qi[k]] for k in range(dim)]) * np.prod(h / 2)
        xq = x0 + (xi + 1) * h / 2
        phi = np.zeros(nd)
        grad = np.zeros((nd, dim))
        for a in range(nd):
            ma = np.unravel_index(a, [p1] * dim, order="F")
            vals = [v1[qi[k], ma[k]] for k in range(dim)]
            phi[a] = np.prod(vals)
            for k in range(dim):
                g = dv1[qi[k], ma[k]] * 2 / h[k]
                for j in range(dim):
                    if j != k:
                        g *= vals[j]
                grad[a, k] = g
        em += w * k_fun(xq) * grad @ grad.T
        ev += w * f_fun(xq) * phi
    return em, ev


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_variable_coefficient_element_against_independent_quadrature(dim, order):
    """K_SIN3 + F_MANUFACTURED element matrices and vectors on an anisotropic box
    (unequal cell counts and lengths per axis) equal an independent numpy
    integration: elmat to 1e-13, elvec to the finite-difference accuracy of the
    source (a wrong axis, sign of grad k . grad u, factor or cell index fails)."""
    n = (4, 3, 5)[:dim]
    lim = [0.0, 2.0, -1.0, 0.5, 0.0, 3.0] if dim == 3 else [0.0, 2.0, -1.0, 0.5, 0.0, 0.0]
    sp = O.Space(O.make_problem(dim, n, order, lim=lim))
    ph = O.make_physics(k_kind=O.K_SIN3, f_kind=O.F_MANUFACTURED)
    for cell in (0, 5, sp.num_cells - 2):
        em, ev = sp.element(ph, cell)
        em_ref, ev_ref = _element_numpy(dim, order, n, lim, cell, _k_sin3, _source_fd)
        assert np.max(np.abs(em - em_ref)) <= 1e-13 * np.max(np.abs(em_ref)), cell
        assert np.max(np.abs(ev - ev_ref)) <= 1e-9 * np.max(np.abs(ev_ref)), cell


@pytest.mark.parametrize("order,n_pair,l2_rate,h1_rate", [(1, (16, 32), 4.0, 2.0), (2, (8, 16), 8.0, 4.0)])
def test_3d_variable_coefficient_converges_at_the_theoretical_rates(order, n_pair, l2_rate, h1_rate):
    """-div(k grad u) = f in 3-D with k = 1 + 0.5 sin sin sin and the manufactured
    source of u = sin sin sin: halving h divides the L2 / H1-seminorm errors by
    2^(p+1) / 2^p within 2 %.  Operator and source must agree on k and on the
    grad k . grad u term: a dropped or sign-flipped term leaves a fixed error
    (ratios near 1)."""
    errs = []
    for n in n_pair:
        sp = O.Space(O.make_problem(3, (n, n, n), order))
        ph = O.make_physics(k_kind=O.K_SIN3, f_kind=O.F_MANUFACTURED)
        irp, ja, val, rhs = sp.assemble(ph, True)
        x, info = O.cg(irp, ja, val, rhs, True, precond="jacobi", rtol=1e-13, maxit=100000)
        assert info["converged"]
        errs.append(sp.error_norms(x))
    a, b = errs
    assert abs(a["l2_norm"] / b["l2_norm"] / l2_rate - 1) < 0.02
    assert abs(a["h1_seminorm"] / b["h1_seminorm"] / h1_rate - 1) < 0.02


def test_cell_coefficient_reads_its_own_cell():
    """K_CELL: each cell's element equals the K_CONST element with k0 = k_cell[c]
    bit for bit, so the per-cell field is read at the cell's own index; the
    assembled quadratic form v^T A v equals sum_c k_c v_c^T E v_c computed from
    the DOF lists (no Dirichlet), and a constant field reproduces K_CONST."""
    n = (5, 4, 3)
    sp = O.Space(O.make_problem(3, n, 1, bc=O.BC_NONE))
    kc = np.random.default_rng(7).uniform(0.1, 10.0, sp.num_cells)
    ph = O.make_physics(k_kind=O.K_CELL, k_cell=kc)
    for c in (0, 1, 7, sp.num_cells - 1):
        em, _ = sp.element(ph, c)
        em_c, _ = sp.element(O.make_physics(k_kind=O.K_CONST, k0=float(kc[c])), c)
        assert np.array_equal(em, em_c), c
    irp, ja, val, rhs = sp.assemble(ph, False)
    v = np.random.default_rng(8).standard_normal(sp.n_free)
    E1, _ = sp.element(O.make_physics(k_kind=O.K_CONST, k0=1.0), 0)   # every cell is a translate
    d = sp.dofs() - 1
    quad = sum(kc[c] * v[d[c]] @ E1 @ v[d[c]] for c in range(sp.num_cells))
    assert abs(v @ O.spmv(irp, ja, val, v, False) - quad) <= 1e-12 * abs(quad)
    const = O.make_physics(k_kind=O.K_CELL, k_cell=np.full(sp.num_cells, 2.5))
    _, _, val_c, _ = sp.assemble(const, False)
    _, _, val_k, _ = sp.assemble(O.make_physics(k_kind=O.K_CONST, k0=2.5), False)
    assert np.array_equal(val_c, val_k)
