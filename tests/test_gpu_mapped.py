"""Parity of the per-cell integration path (moved vertices, coefficients given
at the quadrature points) with the oracle, through the C-ABI.

Gates as for the uniform path: numbering and CSR pattern identical, Dirichlet
values within 1e-14, values and RHS within 1e-12 relative, CG solution within
1e-10 relative of the oracle's FP64 solve (FP32-refined: 1e-8), error norms
within 1e-10; in addition the assembled matrix is bitwise symmetric, the
linear patch test is exact on a moved-vertex mesh, and the manufactured
solution converges at the textbook rates on a smoothly mapped mesh.
"""
import numpy as np
import pytest

import fempar_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

VAL_RTOL = 1e-12
SOL_RTOL = 1e-10


def uniform_vertices(dim, n, lim):
    axes = [(lim[2 * k + 1] - lim[2 * k]) * (np.arange(n[k] + 1) / n[k]) + lim[2 * k] for k in range(dim)]
    grids = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.transpose().ravel() for g in grids], axis=1)


def jittered(dim, n, lim, amp=0.25, seed=3):
    v = uniform_vertices(dim, n, lim)
    idx = np.meshgrid(*[np.arange(n[k] + 1) for k in range(dim)], indexing="ij")
    inner = np.ones(idx[0].shape, dtype=bool)
    for k in range(dim):
        inner &= (idx[k] > 0) & (idx[k] < n[k])
    inner = inner.transpose().ravel()
    h = np.array([(lim[2 * k + 1] - lim[2 * k]) / n[k] for k in range(dim)])
    d = np.random.default_rng(seed).uniform(-amp, amp, size=v.shape) * h
    v[inner] += d[inner]
    return v


def smooth(dim, n, eps=0.08):
    v = uniform_vertices(dim, n, [0, 1, 0, 1, 0, 1])
    bump = np.prod(np.sin(np.pi * v), axis=1)
    w = v.copy()
    for k in range(dim):
        w[:, k] = v[:, k] + eps * np.sin(2 * np.pi * v[:, (k + 1) % dim]) * bump
    return w


def kfun(x):
    return np.exp(0.8 * np.sin(3.0 * x[..., 0]) + x[..., 1] * x[..., -1])


def ffun(x):
    return 1.0 + x[..., 0] ** 2 - 0.5 * x[..., 1]


CASES = {
    # name: (dim, n, order, ncomp, bc, vertices, physics)
    "2d_q1_jit": (2, (24, 20), 1, 1, "boundary", "jit", dict(k="sin3", f="manufactured")),
    "2d_q2_jit": (2, (10, 8), 2, 1, "boundary", "jit", dict(k="sin3", f="manufactured")),
    "3d_q1_jit": (3, (10, 8, 9), 1, 1, "boundary", "jit", dict(k="sin3", f="manufactured")),
    "3d_q2_jit": (3, (5, 4, 6), 2, 1, "boundary", "jit", dict(k="sin3", f="manufactured")),
    "3d_q1_x0_cellk": (3, (9, 7, 6), 1, 1, "x0", "jit", dict(k="cell", f="cell_sin")),
    "3d_q1_qpoint": (3, (8, 9, 7), 1, 1, "boundary", "jit", dict(k=kfun, f=ffun)),
    "3d_q2_qpoint_uniform": (3, (4, 5, 4), 2, 1, "boundary", None, dict(k=kfun, f=ffun)),
    "2d_q2_qpoint_f": (2, (9, 11), 2, 1, "boundary", "jit", dict(k="const", f=ffun)),
    "elasticity_jit": (3, (6, 5, 4), 1, 3, "x0", "jit", dict(elasticity=True, E=40.0, nu=0.3)),
}


def _build(env, name, symmetric_storage=True):
    dim, n, order, ncomp, bc, verts, kw = CASES[name]
    lim = [0, 1, 0, 1, 0, 1] if dim == 3 else [0, 1, 0, 1, 0, 0]
    vc = jittered(dim, n, lim) if verts == "jit" else None
    ncells = int(np.prod(n))
    kc = np.exp(np.random.default_rng(7).normal(0.0, 1.0, ncells)) if kw.get("k") == "cell" else None
    osp = O.Space(O.make_problem(dim, n, order, ncomp, {"boundary": O.BC_ALL, "x0": O.BC_X0}[bc], lim=list(lim),
                                 vcoords=vc))
    if kw.get("elasticity"):
        oph = O.make_physics(kind=O.PHYS_ELASTICITY, E=kw["E"], nu=kw["nu"], body=(0, 0, -1))
        di = F.linear_elasticity_discrete_integration_t(E=kw["E"], nu=kw["nu"], body=(0, 0, -1))
    else:
        xq = osp.quadrature_points()
        k, f = kw["k"], kw["f"]
        args = {}
        if callable(k):
            args.update(k_kind=O.K_QPOINT, k_qp=k(xq))
        else:
            args.update(k_kind={"const": O.K_CONST, "sin3": O.K_SIN3, "cell": O.K_CELL}[k], k0=1.0, k_cell=kc)
        if callable(f):
            args.update(f_kind=O.F_QPOINT, f_qp=f(xq))
        else:
            args.update(f_kind={"manufactured": O.F_MANUFACTURED, "cell_sin": O.F_CELL_SIN}[f], f0=1.0)
        oph = O.make_physics(**args)
        di = F.poisson_discrete_integration_t(k=k, k_cell=kc, f=f, f0=1.0)
    tria = F.serial_triangulation_t().create(env, {
        "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": dim, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": list(n),
        "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": list(lim)}, vertex_coordinates=vc)
    fes = F.serial_fe_space_t().create(tria, order=order, num_components=ncomp, dirichlet=bc)
    op = F.fe_affine_operator_t()
    op.create("CSR", [symmetric_storage], fes, di)
    return osp, oph, fes, op, di


@pytest.mark.parametrize("name", list(CASES))
def test_mapped_pattern_values_rhs(env, name):
    print(f"[mapped] {name}", flush=True)
    osp, oph, fes, op, di = _build(env, name)
    assert np.array_equal(fes.get_lst_dofs_gids(), osp.dofs())
    xq = fes.get_quadrature_points_coordinates()
    assert np.max(np.abs(xq - osp.quadrature_points())) <= 1e-14
    op.compute()
    irp, ja, val, rhs = osp.assemble(oph, True)
    A = op.get_matrix()
    assert np.array_equal(A.irp, irp) and np.array_equal(A.ja, ja)
    assert np.max(np.abs(A.val - val)) <= VAL_RTOL * np.max(np.abs(val))
    b = op.get_translation()
    assert np.max(np.abs(b - rhs)) <= VAL_RTOL * max(np.max(np.abs(rhs)), 1e-300)


@pytest.mark.parametrize("name", ["3d_q1_jit", "3d_q2_jit", "elasticity_jit", "3d_q1_qpoint"])
def test_mapped_matrix_is_bitwise_symmetric(env, name):
    osp, oph, fes, op, di = _build(env, name, symmetric_storage=False)
    op.compute()
    A = op.get_matrix()
    n = len(A.irp) - 1
    rows = np.repeat(np.arange(n), np.diff(A.irp))
    cols = A.ja - 1
    key = rows.astype(np.int64) * n + cols
    tkey = cols.astype(np.int64) * n + rows
    order, torder = np.argsort(key), np.argsort(tkey)
    assert np.array_equal(key[order], tkey[torder])
    assert np.array_equal(A.val[order], A.val[torder])


@pytest.mark.parametrize("name,precision", [("2d_q1_jit", "FP64"), ("3d_q1_jit", "FP64"), ("3d_q2_jit", "FP64"),
                                            ("3d_q1_x0_cellk", "FP64"), ("3d_q1_qpoint", "FP64"),
                                            ("elasticity_jit", "FP64"), ("3d_q1_jit", "FP32_REFINED"),
                                            ("3d_q2_jit", "FP32_REFINED"), ("3d_q1_qpoint", "FP32_REFINED")])
def test_mapped_cg_matches_oracle(env, name, precision):
    print(f"[mapped cg] {name} {precision}", flush=True)
    osp, oph, fes, op, di = _build(env, name)
    op.compute()
    irp, ja, val, rhs = osp.assemble(oph, True)
    x_ref, info = O.cg(irp, ja, val, rhs, True, precond="jacobi", rtol=1e-10, maxit=20000)
    solver = F.iterative_linear_solver_t().create(env)
    solver.set_type_from_string("CG")
    solver.set_parameters_from_pl({"ILS_RTOL": 1e-10, "ILS_MAX_NUM_ITERATIONS": 20000,
                                   "FEMHIP_PRECISION": precision})
    solver.set_operators(op.get_tangent(), F.jacobi_preconditioner_t(op))
    x = solver.solve(op.get_translation())
    r = solver.result
    assert r["did_converge"] and info["converged"]
    tol = SOL_RTOL if precision == "FP64" else 1e-8
    assert np.linalg.norm(x - x_ref) <= tol * np.linalg.norm(x_ref)
    if precision == "FP64":
        assert abs(r["num_iterations"] - info["iterations"]) <= max(2, info["iterations"] // 50)
    Ax = O.spmv(irp, ja, val, x, True)
    assert np.linalg.norm(rhs - Ax) <= 1e-10 * np.linalg.norm(rhs)


@pytest.mark.parametrize("dim,n,order", [(2, (12, 9), 2), (3, (6, 5, 7), 1), (3, (3, 4, 3), 2)])
def test_linear_patch_test_on_the_device(env, dim, n, order):
    """u = x + y (+ z) is reproduced exactly on a moved-vertex mesh (the
    reference's test_poisson check): the nodal values of the device solution
    equal the oracle's interpolation of u at the mapped nodes."""
    lim = [0, 1, 0, 1, 0, 1] if dim == 3 else [0, 1, 0, 1, 0, 0]
    vc = jittered(dim, n, lim, seed=17)
    osp = O.Space(O.make_problem(dim, n, order, lim=list(lim), vcoords=vc))
    oph = O.make_physics(k_kind=O.K_CONST, k0=1.0, f_kind=O.F_CONST, f0=0.0, g_kind=O.G_LINEAR)
    g = osp.fixed_values(oph)
    tria = F.serial_triangulation_t().create(env, {
        "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": dim, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": list(n),
        "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": list(lim)}, vertex_coordinates=vc)
    fes = F.serial_fe_space_t().create(tria, order=order)
    di = F.poisson_discrete_integration_t(k="const", f="const", f0=0.0, g="linear")
    g_hip = fes.interpolate_dirichlet_values(di)
    assert np.max(np.abs(g_hip - g)) <= 1e-14
    op = F.fe_affine_operator_t()
    op.create("CSR", [True], fes, di)
    op.set_fixed_dof_values(g_hip)
    op.compute()
    solver = F.iterative_linear_solver_t().create(env)
    solver.set_type_from_string("CG")
    solver.set_parameters_from_pl({"ILS_RTOL": 1e-14, "ILS_MAX_NUM_ITERATIONS": 20000})
    solver.set_operators(op.get_tangent(), F.jacobi_preconditioner_t(op))
    solver.solve(op.get_translation())
    norms = solver.error_norms("linear")
    assert norms["linfty_norm"] < 1e-10 and norms["h1_seminorm"] < 1e-9


@pytest.mark.parametrize("name", ["2d_q1_jit", "2d_q2_jit", "3d_q1_jit", "3d_q2_jit"])
def test_mapped_error_norms_match_oracle(env, name):
    osp, oph, fes, op, di = _build(env, name)
    op.compute()
    solver = F.iterative_linear_solver_t().create(env)
    solver.set_type_from_string("CG")
    solver.set_parameters_from_pl({"ILS_RTOL": 1e-12, "ILS_MAX_NUM_ITERATIONS": 20000})
    solver.set_operators(op.get_tangent(), F.jacobi_preconditioner_t(op))
    x = solver.solve(op.get_translation())
    got = solver.error_norms("sin")
    ref = osp.error_norms(x, None, exact="sin")
    for key in ("l2_norm", "h1_seminorm", "h1_norm", "linfty_norm"):
        assert abs(got[key] - ref[key]) <= 1e-10 * ref[key], (key, got[key], ref[key])


@pytest.mark.parametrize("dim,order,ns", [(3, 1, (16, 32)), (3, 2, (8, 16))])
def test_mapped_convergence_on_the_device(env, dim, order, ns):
    """Manufactured u on a smoothly mapped cube, solved and measured on the
    device: L2 / H1 errors fall by 4 / 2 (Q1) and 8 / 4 (Q2) per halving of h."""
    errs = []
    for n in ns:
        nn = (n,) * dim
        tria = F.serial_triangulation_t().create(env, {
            "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": dim, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": list(nn),
            "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": [0, 1, 0, 1, 0, 1]}, vertex_coordinates=smooth(dim, nn))
        fes = F.serial_fe_space_t().create(tria, order=order)
        op = F.fe_affine_operator_t()
        op.create("CSR", [True], fes, F.poisson_discrete_integration_t(k="sin3", f="manufactured"))
        op.compute()
        solver = F.iterative_linear_solver_t().create(env)
        solver.set_type_from_string("CG")
        solver.set_parameters_from_pl({"ILS_RTOL": 1e-12, "ILS_MAX_NUM_ITERATIONS": 20000})
        solver.set_operators(op.get_tangent(), F.jacobi_preconditioner_t(op))
        solver.solve_resident()
        e = solver.error_norms("sin")
        errs.append((e["l2_norm"], e["h1_seminorm"]))
    rl2, rh1 = errs[0][0] / errs[1][0], errs[0][1] / errs[1][1]
    want = (4.0, 2.0) if order == 1 else (8.0, 4.0)
    print(f"[mapped rates] Q{order}: L2 {rl2:.3f} H1 {rh1:.3f}", flush=True)
    assert abs(rl2 - want[0]) <= 0.1 * want[0] and abs(rh1 - want[1]) <= 0.1 * want[1]
