"""CPU pins of the oracle's per-cell geometry (moved vertices) and of
coefficients given at the quadrature points.

Each check is one that a plausible mistake in those branches fails:
- the per-cell map applied to the generator's own vertices must reproduce the
  (separately pinned) first-cell-map path;
- element row sums vanish (constants are in the kernel of the Laplacian) and
  the energy of a linear field a.x is |a|^2 times the cell volume, summed over
  a moved-vertex mesh of the unit cube to |a|^2 (a transposed inverse
  Jacobian or a wrong determinant breaks it);
- the f = 1 element vectors sum to the domain's volume;
- the discrete solution reproduces u = x + y (+ z) exactly on a moved-vertex
  mesh (the reference's test_poisson check, test_poisson_driver.f90:562-600),
  which needs the Dirichlet nodes at their mapped positions;
- point-valued k and f reproduce the closed-form kinds at the same points;
- on a smoothly mapped mesh sequence the manufactured solution converges at
  the textbook rates (L2 4 / H1 2 for Q1, 8 / 4 for Q2 per halving of h).
"""
import numpy as np
import pytest

from oracle import oracle as O


def uniform_vertices(dim, n):
    axes = [np.arange(n[k] + 1) / n[k] for k in range(dim)]
    grids = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.transpose().ravel() for g in grids], axis=1)


def interior_mask(dim, n):
    idx = np.meshgrid(*[np.arange(n[k] + 1) for k in range(dim)], indexing="ij")
    m = np.ones(idx[0].shape, dtype=bool)
    for k in range(dim):
        m &= (idx[k] > 0) & (idx[k] < n[k])
    return m.transpose().ravel()


def jittered_vertices(dim, n, amp=0.25, seed=3):
    """Interior vertices moved by up to amp of a cell width; boundary vertices
    stay, so the domain is still the unit square / cube."""
    v = uniform_vertices(dim, n)
    rng = np.random.default_rng(seed)
    h = np.array([1.0 / n[k] for k in range(dim)])
    d = rng.uniform(-amp, amp, size=v.shape) * h
    v[interior_mask(dim, n)] += d[interior_mask(dim, n)]
    return v


def smooth_map(dim, n, eps=0.08):
    """x + eps sin(2 pi x) prod_k sin(pi x_k) along every axis: smooth, the identity on the boundary."""
    v = uniform_vertices(dim, n)
    bump = np.prod(np.sin(np.pi * v), axis=1)
    w = v.copy()
    for k in range(dim):
        w[:, k] = v[:, k] + eps * np.sin(2 * np.pi * v[:, (k + 1) % dim]) * bump
    return w


def node_positions(space, vc):
    """Physical position of every (cell, local DOF) node: the trilinear map of
    its reference position xi = (2/order) m - 1."""
    p = space.problem
    dim, order = p.dim, p.order
    nn = (order + 1) ** dim
    nv = [p.n[k] + 1 for k in range(dim)]
    out = np.zeros((space.num_cells, nn, dim))
    for c in range(space.num_cells):
        ijk = []
        r = c
        for k in range(dim):
            ijk.append(r % p.n[k])
            r //= p.n[k]
        xv = []
        for a in range(1 << dim):
            vid, mul = 0, 1
            for k in range(dim):
                vid += mul * (ijk[k] + ((a >> k) & 1))
                mul *= nv[k]
            xv.append(vc[vid])
        for l in range(nn):
            rem, xi = l, []
            for k in range(dim):
                xi.append((2.0 / order) * (rem % (order + 1)) - 1.0)
                rem //= order + 1
            x = np.zeros(dim)
            for a in range(1 << dim):
                N = 1.0
                for k in range(dim):
                    N *= 0.5 * (1 + xi[k]) if (a >> k) & 1 else 0.5 * (1 - xi[k])
                x += N * xv[a]
            out[c, l] = x
    return out


@pytest.mark.parametrize("dim,n,order", [(2, (5, 4), 1), (2, (3, 3), 2), (3, (3, 2, 4), 1), (3, (2, 2, 2), 2)])
def test_own_map_on_generator_vertices_equals_translated_path(dim, n, order):
    lim = [0.0, 2.0, -1.0, 0.5, 0.0, 3.0] if dim == 3 else [0.0, 2.0, -1.0, 0.5, 0.0, 0.0]
    su = O.Space(O.make_problem(dim, n, order, lim=lim))
    axes = [(lim[2 * k + 1] - lim[2 * k]) * (np.arange(n[k] + 1) / n[k]) + lim[2 * k] for k in range(dim)]
    grids = np.meshgrid(*axes, indexing="ij")
    vc = np.stack([g.transpose().ravel() for g in grids], axis=1)
    sm = O.Space(O.make_problem(dim, n, order, lim=lim, vcoords=vc))
    ph = O.make_physics(k_kind=O.K_SIN3, f_kind=O.F_MANUFACTURED)
    for c in range(su.num_cells):
        eu, vu = su.element(ph, c)
        em, vm = sm.element(ph, c)
        assert np.max(np.abs(eu - em)) <= 1e-13 * np.max(np.abs(eu))
        assert np.max(np.abs(vu - vm)) <= 1e-13 * max(1e-300, np.max(np.abs(vu)))


@pytest.mark.parametrize("dim,n,order", [(2, (4, 5), 1), (2, (3, 4), 2), (3, (3, 3, 3), 1), (3, (2, 3, 2), 2)])
def test_moved_vertices_kernel_energy_and_volume(dim, n, order):
    vc = jittered_vertices(dim, n)
    s = O.Space(O.make_problem(dim, n, order, bc=O.BC_NONE, vcoords=vc))
    ph = O.make_physics(k_kind=O.K_CONST, k0=1.0, f_kind=O.F_CONST, f0=1.0)
    pos = node_positions(s, vc)
    a = np.array([0.7, -1.3, 0.4][:dim])
    energy, vol = 0.0, 0.0
    for c in range(s.num_cells):
        em, ev = s.element(ph, c)
        scale = np.max(np.abs(em))
        assert np.max(np.abs(em.sum(axis=1))) <= 1e-12 * scale      # K 1 = 0
        assert np.array_equal(em, em.T) or np.max(np.abs(em - em.T)) <= 1e-14 * scale
        u = pos[c] @ a                                              # nodal values of a . x
        energy += u @ em @ u
        vol += ev.sum()                                             # sum_i int phi_i = |cell|
    assert abs(vol - 1.0) <= 1e-12
    assert abs(energy - a @ a) <= 1e-11 * (a @ a)


def test_energy_detects_a_transposed_inverse_jacobian():
    """The energy pin is sensitive: on a sheared mesh, using J^-1 instead of
    J^-T for the physical gradients changes a.x's energy."""
    dim, n = 2, (3, 3)
    v = uniform_vertices(dim, n)
    v[:, 0] += 0.3 * v[:, 1]                      # affine shear: every cell a parallelogram
    s = O.Space(O.make_problem(dim, n, 1, bc=O.BC_NONE, vcoords=v))
    qp, qw, shp, dshp = s.tables()
    pos = node_positions(s, v)
    a = np.array([1.0, 0.5])
    J = np.array([[1.0 / 3 / 2, 0.3 / 3 / 2], [0.0, 1.0 / 3 / 2]])    # d x / d xi of a cell
    good = bad = 0.0
    for c in range(s.num_cells):
        u = pos[c] @ a
        for q in range(s.nq):
            gref = dshp[q] .T @ u                  # reference gradient of u_h
            good += np.dot(np.linalg.inv(J).T @ gref, np.linalg.inv(J).T @ gref) * abs(np.linalg.det(J)) * qw[q]
            bad += np.dot(np.linalg.inv(J) @ gref, np.linalg.inv(J) @ gref) * abs(np.linalg.det(J)) * qw[q]
    ph = O.make_physics(k_kind=O.K_CONST, k0=1.0)
    energy = sum(pos[c] @ a @ s.element(ph, c)[0] @ (pos[c] @ a) for c in range(s.num_cells))
    vol = 1.0   # a shear keeps the unit square's area
    assert abs(good - (a @ a) * vol) < 1e-12 and abs(bad - good) > 1e-2
    assert abs(energy - good) < 1e-12


@pytest.mark.parametrize("dim,n,order", [(2, (6, 5), 1), (2, (4, 3), 2), (3, (4, 3, 3), 1), (3, (2, 3, 2), 2)])
def test_linear_patch_test_on_moved_vertices(dim, n, order):
    vc = jittered_vertices(dim, n, seed=11)
    s = O.Space(O.make_problem(dim, n, order, vcoords=vc))
    ph = O.make_physics(k_kind=O.K_CONST, k0=1.0, f_kind=O.F_CONST, f0=0.0, g_kind=O.G_LINEAR)
    fx = s.fixed_values(ph)
    irp, ja, val, rhs = s.assemble(ph, True, fixed_values=fx)
    x, info = O.cg(irp, ja, val, rhs, True, precond="jacobi", rtol=1e-14)
    dofs = s.dofs()
    pos = node_positions(s, vc)
    exact = np.zeros(s.n_free)
    for c in range(s.num_cells):
        for l in range(s.nd):
            if dofs[c, l] > 0:
                exact[dofs[c, l] - 1] = pos[c, l].sum()
    assert np.max(np.abs(x - exact)) <= 1e-10
    # nodal interpolation of g at the mapped boundary nodes
    for c in range(s.num_cells):
        for l in range(s.nd):
            if dofs[c, l] < 0:
                assert abs(fx[-dofs[c, l] - 1] - pos[c, l].sum()) <= 1e-14


@pytest.mark.parametrize("dim,order", [(2, 1), (3, 2)])
def test_point_values_reproduce_closed_form_kinds(dim, order):
    n = (3, 4, 2)[:dim]
    vc = jittered_vertices(dim, n, seed=5)
    s = O.Space(O.make_problem(dim, n, order, vcoords=vc))
    xq = s.quadrature_points()
    # k = 2.5 at every point scales the k = 1 matrix exactly
    one = O.make_physics(k_kind=O.K_CONST, k0=1.0)
    qk = O.make_physics(k_kind=O.K_QPOINT, k_qp=np.full(xq.shape[:2], 2.5))
    for c in range(s.num_cells):
        e1, _ = s.element(one, c)
        e2, _ = s.element(qk, c)
        assert np.max(np.abs(e2 - 2.5 * e1)) <= 1e-14 * np.max(np.abs(e2))
    # k and f at the points from the closed forms give the closed-form kinds' elements
    ks = 1.0 + 0.5 * np.prod(np.sin(2 * np.pi * xq), axis=2)
    ref = O.make_physics(k_kind=O.K_SIN3, f_kind=O.F_CONST, f0=0.0)
    pts = O.make_physics(k_kind=O.K_QPOINT, k_qp=ks, f_kind=O.F_CONST, f0=0.0)
    for c in range(s.num_cells):
        e1, _ = s.element(ref, c)
        e2, _ = s.element(pts, c)
        assert np.max(np.abs(e2 - e1)) <= 1e-13 * np.max(np.abs(e1))
    # the manufactured source at the points gives the manufactured kind's vector
    d = dim
    s1, c1 = np.sin(np.pi * xq), np.cos(np.pi * xq)
    s2, c2 = np.sin(2 * np.pi * xq), np.cos(2 * np.pi * xq)
    u = np.prod(s1, axis=2)
    gk = np.stack([np.pi * c2[..., j] * np.prod(np.delete(s2, j, axis=2), axis=2) for j in range(d)], axis=2)
    gu = np.stack([np.pi * c1[..., j] * np.prod(np.delete(s1, j, axis=2), axis=2) for j in range(d)], axis=2)
    fm = ks * d * np.pi ** 2 * u - np.sum(gk * gu, axis=2)
    ref = O.make_physics(k_kind=O.K_SIN3, f_kind=O.F_MANUFACTURED)
    pts = O.make_physics(k_kind=O.K_QPOINT, k_qp=ks, f_kind=O.F_QPOINT, f_qp=fm)
    one_f = O.make_physics(k_kind=O.K_CONST, k0=1.0, f_kind=O.F_CONST, f0=1.0)
    unit_pts = O.make_physics(k_kind=O.K_CONST, k0=1.0, f_kind=O.F_QPOINT, f_qp=np.ones(xq.shape[:2]))
    for c in range(s.num_cells):
        _, v1 = s.element(ref, c)
        _, v2 = s.element(pts, c)
        assert np.max(np.abs(v2 - v1)) <= 1e-12 * np.max(np.abs(v1))
        _, w1 = s.element(one_f, c)
        _, w2 = s.element(unit_pts, c)
        assert np.array_equal(w1, w2)


def _h_errors(dim, order, n, eps):
    nn = (n,) * dim
    vc = smooth_map(dim, nn, eps)
    s = O.Space(O.make_problem(dim, nn, order, vcoords=vc))
    ph = O.make_physics(k_kind=O.K_SIN3, f_kind=O.F_MANUFACTURED)
    irp, ja, val, rhs = s.assemble(ph, True)
    x, info = O.cg(irp, ja, val, rhs, True, precond="jacobi", rtol=1e-13, maxit=20000)
    e = s.error_norms(x)
    return e["l2_norm"], e["h1_seminorm"]


@pytest.mark.parametrize("dim,order,ns,rates", [(2, 1, (16, 32, 64), (4.0, 2.0)), (2, 2, (8, 16, 32), (8.0, 4.0)),
                                                (3, 1, (8, 16), (4.0, 2.0))])
def test_convergence_rates_on_a_smoothly_mapped_mesh(dim, order, ns, rates):
    errs = [_h_errors(dim, order, n, 0.08) for n in ns]
    for (l2a, h1a), (l2b, h1b) in zip(errs, errs[1:]):
        assert abs(l2a / l2b - rates[0]) <= 0.08 * rates[0], (l2a / l2b, rates)
        assert abs(h1a / h1b - rates[1]) <= 0.08 * rates[1], (h1a / h1b, rates)
