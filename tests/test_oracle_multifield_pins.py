"""Pins of the oracle's multi-field restatement (several scalar fields on one
mesh, block layout, block assembly).  CPU only.

Numbering offsets are checked against hand-derived ids; the element against
the single-field element (sigma = 0) and against an independent exact
integration of the mixed mass matrices (kappa = 0); the block assembly by
consistency between layouts; the whole chain by reduction to the single-field
solve when both fields carry the same problem.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _q(dim, n, order, bc):
    return O.make_problem(dim, n, order, bc=bc)


def test_offsets_on_a_2x2_mesh():
    """u: Q1 without Dirichlet (9 free), v: Q1 fixed on the boundary (1 free,
    8 fixed).  One block: v's free id follows u's 9; two blocks: v restarts
    at 1.  Fixed ids follow the earlier fields' fixed ids."""
    u, v = _q(2, (2, 2), 1, O.BC_NONE), _q(2, (2, 2), 1, O.BC_ALL)
    one = O.MultiFieldSpace([u, v], [1, 1])
    assert one.block_sizes == [10] and one.n_free == [9, 1] and one.n_fixed == [0, 8]
    assert one.dofs(0).tolist() == [[1, 2, 3, 4], [2, 5, 4, 6], [3, 4, 7, 8], [4, 6, 8, 9]]
    assert one.dofs(1).tolist() == [[-1, -2, -3, 10], [-2, -4, 10, -5], [-3, 10, -6, -7], [10, -5, -7, -8]]
    two = O.MultiFieldSpace([u, v], [1, 2])
    assert two.block_sizes == [9, 1]
    assert (two.dofs(1) == np.where(one.dofs(1) > 0, 1, one.dofs(1))).all()
    both = O.MultiFieldSpace([v, v], [1, 1])          # both fixed on the boundary
    assert both.dofs(0)[0].tolist() == [-1, -2, -3, 1] and both.dofs(1)[0].tolist() == [-9, -10, -11, 2]


@pytest.mark.parametrize("dim,orders", [(2, (2, 1)), (3, (1, 2)), (2, (1, 1))])
def test_each_field_numbers_as_if_alone(dim, orders):
    n = (3, 4) if dim == 2 else (2, 3, 2)
    probs = [_q(dim, n, o, O.BC_ALL) for o in orders]
    mf = O.MultiFieldSpace(probs, [1, 1])
    off = 0
    for f, p in enumerate(probs):
        single = O.Space(p).dofs()
        got = mf.dofs(f)
        assert np.array_equal(got[single > 0], single[single > 0] + off)
        off += O.Space(p).n_free


@pytest.mark.parametrize("dim,orders", [(2, (2, 1)), (3, (1, 2))])
def test_element_without_coupling_is_the_poisson_element(dim, orders):
    n = (4, 3) if dim == 2 else (2, 2, 3)
    probs = [_q(dim, n, o, O.BC_ALL) for o in orders]
    mf = O.MultiFieldSpace(probs, [1, 1])
    kappa = (1.5, 0.25)
    em, ev = mf.element(O.make_mf_physics(kappa, 0.0, (2.0, -1.0)))
    nd = [O.Space(p).nd for p in probs]
    off = [0, nd[0]]
    for f, p in enumerate(probs):
        sp = O.Space(p)
        ref, rv = sp.element(O.make_physics(k_kind=O.K_CONST, k0=kappa[f], f_kind=O.F_CONST,
                                            f0=(2.0, -1.0)[f]), 0)
        blk = em[off[f]:off[f] + nd[f], off[f]:off[f] + nd[f]]
        assert np.max(np.abs(blk - ref)) <= 1e-13 * np.max(np.abs(ref))
        assert np.max(np.abs(ev[off[f]:off[f] + nd[f]] - rv)) <= 1e-13 * np.max(np.abs(rv))
    assert np.all(em[:nd[0], nd[0]:] == 0) and np.all(em[nd[0]:, :nd[0]] == 0)


def _lagrange_1d(order, h):
    """1-D Lagrange basis on [0, h] through equispaced nodes, as numpy polynomials."""
    nodes = np.linspace(0, h, order + 1)
    basis = []
    for a in range(order + 1):
        others = [nodes[b] for b in range(order + 1) if b != a]
        pol = np.polynomial.Polynomial.fromroots(others)
        basis.append(pol / pol(nodes[a]))
    return basis


def _mass_1d(oa, ob, h):
    x, w = np.polynomial.legendre.leggauss(12)
    xx = 0.5 * h * (x + 1)
    A, B = _lagrange_1d(oa, h), _lagrange_1d(ob, h)
    return np.array([[0.5 * h * np.sum(w * a(xx) * b(xx)) for b in B] for a in A])


def test_coupling_blocks_are_exact_mixed_mass_matrices():
    """kappa = 0, sigma = 1, 2-D Q2 x Q1 on a 3 x 2 box of [0, 1.5] x [0, 1]:
    the (1, 2) block is -(mixed mass matrix), the Kronecker product of 1-D
    mixed mass matrices integrated here with 12-point Gauss on numpy
    polynomials; the element is symmetric; its row sums are the integrals of
    the basis functions (partition of unity), the total is the cell area."""
    lim = [0, 1.5, 0, 1, 0, 0]
    p2 = O.make_problem(2, (3, 2), 2, bc=O.BC_ALL, lim=lim)
    p1 = O.make_problem(2, (3, 2), 1, bc=O.BC_ALL, lim=lim)
    mf = O.MultiFieldSpace([p2, p1], [1, 2])
    em, _ = mf.element(O.make_mf_physics((0.0, 0.0), 1.0, (0.0, 0.0)))
    hx, hy = 0.5, 0.5
    mx, my = _mass_1d(2, 1, hx), _mass_1d(2, 1, hy)
    # node index a = ax + (p+1) ay (x fastest), both fields
    ref = np.zeros((9, 4))
    for a in range(9):
        for b in range(4):
            ref[a, b] = mx[a % 3, b % 2] * my[a // 3, b // 2]
    assert np.max(np.abs(em[:9, 9:] + ref)) <= 1e-14
    assert np.max(np.abs(em - em.T)) <= 1e-15
    m22 = em[9:, 9:]
    assert abs(np.sum(m22) - hx * hy) <= 1e-14              # (nf-1) sigma M22, nf = 2
    assert np.max(np.abs(-em[:9, 9:].sum(axis=1) - em[:9, :9].sum(axis=1))) <= 1e-14
    assert np.max(np.abs(-em[:9, 9:].sum(axis=0) - m22.sum(axis=1))) <= 1e-14


def test_block_layouts_agree():
    """The blocks of the two-block layout are the sub-matrices of the one-block
    layout's matrix (rows and columns shifted by the first field's size), bit
    for bit; without coupling the off-diagonal blocks are empty and the
    one-block matrix has no cross entries."""
    probs = [_q(2, (5, 4), 2, O.BC_ALL), _q(2, (5, 4), 1, O.BC_X0)]
    ph = O.make_mf_physics((1.0, 2.0), 0.75, (1.0, 0.5))
    one = O.MultiFieldSpace(probs, [1, 1])
    two = O.MultiFieldSpace(probs, [1, 2])
    n1 = one.n_free[0]
    full = [[1, 1], [1, 1]]
    irp, ja, val = one.assemble_block(ph, full, 1, 1)
    rows = [(ja[irp[r] - 1:irp[r + 1] - 1], val[irp[r] - 1:irp[r + 1] - 1]) for r in range(len(irp) - 1)]
    for ib in (1, 2):
        for jb in (1, 2):
            bi, bj, bv = two.assemble_block(ph, full, ib, jb)
            for r in range(len(bi) - 1):
                cols, vals = rows[r + (n1 if ib == 2 else 0)]
                sel = (cols > n1) if jb == 2 else (cols <= n1)
                assert np.array_equal(bj[bi[r] - 1:bi[r + 1] - 1] + (n1 if jb == 2 else 0), cols[sel])
                assert np.array_equal(bv[bi[r] - 1:bi[r + 1] - 1], vals[sel])
    rhs1 = one.assemble_rhs(ph, 1)
    assert np.array_equal(np.concatenate([two.assemble_rhs(ph, 1), two.assemble_rhs(ph, 2)]), rhs1)
    diag_only = [[1, 0], [0, 1]]
    bi, bj, bv = two.assemble_block(ph, diag_only, 1, 2)
    assert len(bj) == 0 and np.all(bi == 1)
    irp, ja, val = one.assemble_block(ph, diag_only, 1, 1)
    for r in range(len(irp) - 1):
        c = ja[irp[r] - 1:irp[r + 1] - 1]
        assert np.all(c <= n1) if r < n1 else np.all(c > n1)


def test_symmetric_storage_of_a_diagonal_block_is_its_upper_triangle():
    probs = [_q(3, (3, 2, 2), 1, O.BC_ALL), _q(3, (3, 2, 2), 2, O.BC_ALL)]
    ph = O.make_mf_physics((1.0, 1.0), 0.5, (1.0, 1.0))
    mf = O.MultiFieldSpace(probs, [1, 1])
    irp, ja, val = mf.assemble_block(ph, [[1, 1], [1, 1]], 1, 1, symmetric_storage=False)
    ui, uj, uv = mf.assemble_block(ph, [[1, 1], [1, 1]], 1, 1, symmetric_storage=True)
    for r in range(len(irp) - 1):
        c = ja[irp[r] - 1:irp[r + 1] - 1]
        keep = c >= r + 1
        assert np.array_equal(uj[ui[r] - 1:ui[r + 1] - 1], c[keep])
        assert np.array_equal(uv[ui[r] - 1:ui[r + 1] - 1], val[irp[r] - 1:irp[r + 1] - 1][keep])


@pytest.mark.parametrize("layout", [[1, 1], [1, 2]])
def test_equal_fields_reduce_to_the_single_field_solve(layout):
    """Two copies of the same Poisson problem coupled by sigma (u_1 - u_2, v_1 -
    v_2): the solution has u_1 = u_2 = the single-field solution, whatever
    sigma, so the block system's solve must return it twice."""
    p = _q(2, (12, 10), 2, O.BC_ALL)
    mf = O.MultiFieldSpace([p, p], layout)
    irp, ja, val, rhs = mf.merged(O.make_mf_physics((1.3, 1.3), 5.0, (1.0, 1.0)), [[1, 1], [1, 1]])
    x, info = O.cg(irp, ja, val, rhs, False, precond="jacobi", rtol=1e-13, maxit=5000)
    assert info["converged"]
    sp = O.Space(p)
    si, sj, sv, sr = sp.assemble(O.make_physics(k_kind=O.K_CONST, k0=1.3, f_kind=O.F_CONST, f0=1.0), False)
    xs, _ = O.cg(si, sj, sv, sr, False, precond="jacobi", rtol=1e-13, maxit=5000)
    n = sp.n_free
    assert np.linalg.norm(x[:n] - xs) <= 1e-10 * np.linalg.norm(xs)
    assert np.linalg.norm(x[n:] - xs) <= 1e-10 * np.linalg.norm(xs)
