"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Gates (bit-exact for integers, tolerances written here for floating point):
  * DOF numbering and CSR pattern (irp, ja) identical to the oracle's;
  * assembled values and RHS within 1e-12 relative (max-norm, scaled by the
    largest reference entry);
  * CG solution within 1e-10 relative (2-norm) of the oracle's FP64 solve.
"""
import numpy as np
import pytest

import fempar_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

VAL_RTOL = 1e-12
SOL_RTOL = 1e-10

CASES = {
    # name: (dim, n, order, ncomp, bc, physics kwargs for both sides)
    "2d_q1_64_sin": (2, (64, 64), 1, 1, "boundary", dict(k="const", f="manufactured")),
    "2d_q1_64_f1": (2, (64, 64), 1, 1, "boundary", dict(k="const", f="const", f0=1.0)),
    "2d_q2_12": (2, (12, 12), 2, 1, "boundary", dict(k="sin3", f="manufactured")),
    "3d_q1_20_sin3": (3, (20, 20, 20), 1, 1, "boundary", dict(k="sin3", f="manufactured")),
    "3d_q1_aniso": (3, (9, 6, 5), 1, 1, "boundary", dict(k="sin3", f="manufactured", lim=[0, 2, -1, 0.5, 0, 3])),
    "3d_q2_8_sin3": (3, (8, 8, 8), 2, 1, "boundary", dict(k="sin3", f="manufactured")),
    "3d_q1_cellk": (3, (12, 12, 12), 1, 1, "boundary", dict(k="cell", f="cell_sin")),
    "3d_q1_linear_g": (3, (6, 6, 6), 1, 1, "boundary", dict(k="const", f="zero", g="linear")),
    "2d_q2_linear_g": (2, (5, 5), 2, 1, "boundary", dict(k="const", f="zero", g="linear")),
    "elasticity_8": (3, (8, 8, 8), 1, 3, "x0", dict(elasticity=True, E=50.0, nu=0.3)),
    "3d_q2_x0_7x5x6": (3, (7, 5, 6), 2, 1, "x0", dict(k="sin3", f="manufactured")),
}


def _lim(dim, kw):
    return kw.get("lim") or ([0, 1, 0, 1, 0, 1] if dim == 3 else [0, 1, 0, 1, 0, 0])


def _k_cell(ncells, seed=7):
    return np.exp(np.random.default_rng(seed).normal(0.0, 1.0, ncells))


def _build(env, name, symmetric_storage=True):
    dim, n, order, ncomp, bc, kw = CASES[name]
    lim = _lim(dim, kw)
    ncells = int(np.prod(n))
    kc = _k_cell(ncells) if kw.get("k") == "cell" else None
    # oracle side
    osp = O.Space(O.make_problem(dim, n, order, ncomp, {"boundary": O.BC_ALL, "x0": O.BC_X0}[bc], lim=list(lim)))
    if kw.get("elasticity"):
        oph = O.make_physics(kind=O.PHYS_ELASTICITY, E=kw["E"], nu=kw["nu"], body=(0, 0, -1))
        di = F.linear_elasticity_discrete_integration_t(E=kw["E"], nu=kw["nu"], body=(0, 0, -1))
    else:
        kk = {"const": O.K_CONST, "sin3": O.K_SIN3, "cell": O.K_CELL}[kw["k"]]
        ff = {"manufactured": O.F_MANUFACTURED, "cell_sin": O.F_CELL_SIN, "zero": O.F_CONST,
              "const": O.F_CONST}[kw["f"]]
        gg = O.G_LINEAR if kw.get("g") == "linear" else O.G_ZERO
        f0 = kw.get("f0", 1.0 if kw["f"] == "cell_sin" else 0.0)
        oph = O.make_physics(k_kind=kk, k0=1.0, k_cell=kc, f_kind=ff, f0=f0, g_kind=gg)
        di = F.poisson_discrete_integration_t(k=kw["k"], k_cell=kc,
                                              f="const" if kw["f"] in ("zero", "const") else kw["f"],
                                              f0=f0, g=kw.get("g", "zero"))
    # HIP side
    tria = F.serial_triangulation_t().create(env, {
        "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": dim, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": list(n),
        "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": list(lim)})
    fes = F.serial_fe_space_t().create(tria, order=order, num_components=ncomp, dirichlet=bc)
    op = F.fe_affine_operator_t()
    op.create("CSR", [symmetric_storage], fes, di)
    return osp, oph, fes, op, di


@pytest.mark.parametrize("name", list(CASES))
def test_numbering_pattern_values_rhs(env, name):
    print(f"[parity] {name}", flush=True)
    osp, oph, fes, op, di = _build(env, name, symmetric_storage=True)
    assert (fes.num_free_dofs, fes.num_fixed_dofs) == (osp.n_free, osp.n_fixed)
    assert np.array_equal(fes.get_lst_dofs_gids(), osp.dofs())
    g = osp.fixed_values(oph)
    g_hip = fes.interpolate_dirichlet_values(di)
    assert np.allclose(g_hip, g, rtol=0, atol=1e-14)
    op.set_fixed_dof_values(g_hip if osp.n_fixed else None)
    op.compute()
    irp, ja, val, rhs = osp.assemble(oph, True, fixed_values=g)
    A = op.get_matrix()
    assert A.symmetric_storage
    assert np.array_equal(A.irp, irp) and np.array_equal(A.ja, ja)          # bit-exact pattern
    scale = np.max(np.abs(val))
    assert np.max(np.abs(A.val - val)) <= VAL_RTOL * scale
    b = op.get_translation()
    bscale = max(np.max(np.abs(rhs)), 1e-300)
    assert np.max(np.abs(b - rhs)) <= VAL_RTOL * bscale


@pytest.mark.parametrize("name", ["2d_q1_64_sin", "3d_q1_aniso", "elasticity_8"])
def test_full_storage_pattern_and_apply(env, name):
    osp, oph, fes, op, di = _build(env, name, symmetric_storage=False)
    op.compute()
    irp, ja, val, rhs = osp.assemble(oph, False)
    A = op.get_matrix()
    assert np.array_equal(A.irp, irp) and np.array_equal(A.ja, ja)
    assert np.max(np.abs(A.val - val)) <= VAL_RTOL * np.max(np.abs(val))
    x = np.random.default_rng(3).standard_normal(osp.n_free)
    y = op.apply(x)
    y_ref = O.spmv(irp, ja, val, x, False)
    assert np.max(np.abs(y - y_ref)) <= 1e-12 * np.max(np.abs(y_ref))


@pytest.mark.parametrize("name,prec", [("2d_q1_64_sin", "jacobi"), ("2d_q1_64_f1", "jacobi"),
                                       ("2d_q1_64_f1", "identity"), ("3d_q1_20_sin3", "jacobi"),
                                       ("3d_q2_8_sin3", "jacobi"), ("3d_q1_cellk", "identity"),
                                       ("elasticity_8", "jacobi"), ("3d_q1_linear_g", "identity")])
def test_cg_solution_matches_oracle(env, name, prec):
    print(f"[cg] {name} {prec}", flush=True)
    osp, oph, fes, op, di = _build(env, name)
    g = osp.fixed_values(oph)
    op.set_fixed_dof_values(g if osp.n_fixed else None)
    op.compute()
    irp, ja, val, rhs = osp.assemble(oph, True, fixed_values=g)
    x_ref, info = O.cg(irp, ja, val, rhs, True, precond=prec, rtol=1e-10, maxit=20000)
    solver = F.iterative_linear_solver_t().create(env)
    solver.set_type_from_string("CG")
    solver.set_parameters_from_pl({"ILS_RTOL": 1e-10, "ILS_MAX_NUM_ITERATIONS": 20000})
    M = F.jacobi_preconditioner_t(op) if prec == "jacobi" else F.identity_operator_t(op)
    solver.set_operators(op.get_tangent(), M)
    x = solver.solve(op.get_translation())
    r = solver.result
    assert r["did_converge"] and info["converged"]
    assert abs(r["num_iterations"] - info["iterations"]) <= max(2, info["iterations"] // 50)
    assert np.linalg.norm(x - x_ref) <= SOL_RTOL * np.linalg.norm(x_ref)
    # device-resident solve (operator RHS, x0 = 0) gives the same answer
    solver.solve_resident()
    assert np.linalg.norm(solver.get_solution() - x_ref) <= SOL_RTOL * np.linalg.norm(x_ref)


@pytest.mark.parametrize("name,exact", [("2d_q1_64_sin", "sin"), ("2d_q2_12", "sin"), ("3d_q1_20_sin3", "sin"),
                                        ("3d_q2_8_sin3", "sin"), ("3d_q1_linear_g", "linear"),
                                        ("2d_q2_linear_g", "linear")])
def test_error_norms_match_oracle(env, name, exact):
    """Device error norms of the device solution equal the oracle's norms of the
    same vector within 1e-10 relative (1e-12 absolute when the FE space holds u)."""
    print(f"[norms] {name} {exact}", flush=True)
    osp, oph, fes, op, di = _build(env, name)
    g = osp.fixed_values(oph)
    op.set_fixed_dof_values(g if osp.n_fixed else None)
    op.compute()
    solver = F.iterative_linear_solver_t().create(env)
    solver.set_type_from_string("CG")
    solver.set_parameters_from_pl({"ILS_RTOL": 1e-12, "ILS_MAX_NUM_ITERATIONS": 20000})
    solver.set_operators(op.get_tangent(), F.jacobi_preconditioner_t(op))
    x = solver.solve(op.get_translation())
    got = solver.error_norms(exact)
    ref = osp.error_norms(x, g if osp.n_fixed else None, exact=exact)
    for key in ("l2_norm", "h1_seminorm", "h1_norm", "linfty_norm"):
        if exact == "linear":
            assert got[key] < 1e-10 and ref[key] < 1e-10, key
        else:
            assert abs(got[key] - ref[key]) <= 1e-10 * ref[key], (key, got[key], ref[key])


CRITERIA_CASES = [("3d_q1_20_sin3", 1e-10), ("3d_q1_cellk", 1e-9)]   # 34-40 iterations under every criterion


@pytest.mark.parametrize("case,rtol", CRITERIA_CASES)
@pytest.mark.parametrize("crit", O.CRITERIA)
def test_stopping_criteria_match_oracle(env, crit, case, rtol):
    """All ten stopping criteria of cg.f90:323-407 on systems that take 34-40
    Jacobi-PCG iterations (no criterion crosses its threshold within 30 % of
    it, so rounding cannot move the exit): the same iteration count as the
    oracle, the same rhs_convergence_test, and the error-estimate history of
    every iteration (RES_* the recurrence residual, *NRMGIVEN the lagged
    ||r||_M, DELTA_* alpha ||p|| with its first-iteration threshold) within
    1e-7 relative."""
    print(f"[criteria] {case} {crit}", flush=True)
    osp, oph, fes, op, di = _build(env, case)
    op.compute()
    irp, ja, val, rhs = osp.assemble(oph, True)
    x_ref, info = O.cg(irp, ja, val, rhs, True, precond="jacobi", criterion=crit, rtol=rtol, maxit=5000,
                       history=True)
    assert info["converged"] and info["iterations"] >= 30
    solver = F.iterative_linear_solver_t().create(env)
    solver.set_type_from_string("CG")
    solver.set_parameters_from_pl({"ILS_RTOL": rtol, "ILS_STOPPING_CRITERIUM": crit})
    solver.set_operators(op, F.jacobi_preconditioner_t(op))
    x = solver.solve(rhs)
    assert solver.did_converge()
    assert solver.get_num_iterations() == info["iterations"], crit
    assert abs(solver.result["rhs_convergence_test"] - info["rhs_test"]) <= 1e-9 * info["rhs_test"]
    hist = solver.get_error_estimate_history_convergence_test()
    assert len(hist) == info["iterations"]
    assert np.max(np.abs(hist - info["history"]) / info["history"]) <= 1e-7, crit
    assert np.linalg.norm(x - x_ref) <= 1e-8 * np.linalg.norm(x_ref)


def test_fp32_refined_reaches_fp64_accuracy(env):
    osp, oph, fes, op, di = _build(env, "3d_q1_20_sin3")
    op.compute()
    irp, ja, val, rhs = osp.assemble(oph, True)
    x_ref, info = O.cg(irp, ja, val, rhs, True, precond="jacobi", rtol=1e-12, maxit=20000)
    solver = F.iterative_linear_solver_t().create(env)
    solver.set_type_from_string("CG")
    solver.set_parameters_from_pl({"ILS_RTOL": 1e-10, "ILS_MAX_NUM_ITERATIONS": 20000,
                                   "FEMHIP_PRECISION": "FP32_REFINED"})
    solver.set_operators(op, F.jacobi_preconditioner_t(op))
    x = solver.solve(rhs)
    r = solver.result
    assert r["did_converge"] and r["outer_iterations"] >= 2
    assert r["true_residual"] <= 1.0001e-10 * np.linalg.norm(rhs)
    assert np.linalg.norm(x - x_ref) <= 1e-8 * np.linalg.norm(x_ref)


@pytest.mark.parametrize("name", ["2d_q1_64_f1", "3d_q1_20_sin3", "3d_q1_cellk"])
def test_fp32_integration_on_matrix_cores(env, name, monkeypatch):
    """The refined solver's FP32 matrix integrated in FP32 on the matrix cores
    with FP32 accumulation (FEMHIP_FP32_INTEGRATION=1, scalar Q1) against the
    default, the FP64 matrix rounded to FP32: both solves reach the oracle's
    FP64 solution within 1e-8 with a true FP64 residual within rtol ||b||.
    The integrated matrix may cost inner iterations (its accumulation error is
    a few FP32 ulps, rounding's half an ulp: measured 137 against 120, 43
    against 40), but at most 30 % more; a wrong lane layout of the MFMA
    result or a wrong neighbour code would move entries between rows or
    diagonals and cost far more or fail to converge."""
    print(f"[fp32-integration] {name}", flush=True)
    osp, oph, fes, op, di = _build(env, name)
    op.compute()
    irp, ja, val, rhs = osp.assemble(oph, True)
    x_ref, info = O.cg(irp, ja, val, rhs, True, precond="jacobi", rtol=1e-12, maxit=20000)
    bn = np.linalg.norm(rhs)
    its = {}
    for mode in ("rounded", "mfma"):
        if mode == "mfma":
            monkeypatch.setenv("FEMHIP_FP32_INTEGRATION", "1")
            osp, oph, fes, op, di = _build(env, name)   # a fresh operator: its FP32 copy is made anew
            op.compute()
        s = F.iterative_linear_solver_t().create(env)
        s.set_type_from_string("CG")
        s.set_parameters_from_pl({"ILS_RTOL": 1e-10, "ILS_MAX_NUM_ITERATIONS": 20000,
                                  "FEMHIP_PRECISION": "FP32_REFINED"})
        s.set_operators(op, F.jacobi_preconditioner_t(op))
        x = s.solve(rhs)
        r = s.result
        assert r["did_converge"] and r["true_residual"] <= 1.0001e-10 * bn, mode
        assert np.linalg.norm(x - x_ref) <= 1e-8 * np.linalg.norm(x_ref), mode
        its[mode] = r["num_iterations"]
        s.free()
    assert its["rounded"] - 2 <= its["mfma"] <= its["rounded"] * 1.3 + 2, its


@pytest.mark.parametrize("name", ["2d_q1_64_f1", "3d_q1_20_sin3", "3d_q1_cellk", "3d_q2_8_sin3", "elasticity_8"])
def test_fp32_strict_reaches_fp64_accuracy(env, name):
    """FEMHIP_PRECISION = FP32_STRICT: the FP32 matrix integrated in FP32 on
    the matrix cores (scalar Q1; rounded from FP64 for Q2) and its rows summed
    in FP32, inside the same FP64 refinement.  Solved on one operator by turns
    with an FP32_REFINED solver, so that each solve must make the operator's
    FP32 copy its own way again (vector Q1 and Q2 run the three-kernel slots
    on the class-blocked layout): both reach the oracle's FP64 solution within
    1e-8 with a true FP64 residual within rtol ||b||, and the strict solve takes
    at most 40 % more inner iterations (FP32 row sums of 7 or 27 terms add a
    few ulps per product; a wrong accumulator or a stale copy of the other
    kind costs far more or fails to converge)."""
    print(f"[fp32-strict] {name}", flush=True)
    osp, oph, fes, op, di = _build(env, name)
    op.compute()
    irp, ja, val, rhs = osp.assemble(oph, True)
    x_ref, info = O.cg(irp, ja, val, rhs, True, precond="jacobi", rtol=1e-12, maxit=20000)
    bn = np.linalg.norm(rhs)
    solvers = {p: _solver(env, op, p) for p in ("FP32_STRICT", "FP32_REFINED")}
    its = {}
    for p in ("FP32_STRICT", "FP32_REFINED", "FP32_STRICT"):
        x = solvers[p].solve(rhs)
        r = solvers[p].result
        assert r["did_converge"] and r["true_residual"] <= 1.0001e-10 * bn, p
        assert np.linalg.norm(x - x_ref) <= 1e-8 * np.linalg.norm(x_ref), p
        assert its.setdefault(p, r["num_iterations"]) == r["num_iterations"], p   # the second strict solve repeats
    for s in solvers.values():
        s.free()
    assert its["FP32_STRICT"] <= its["FP32_REFINED"] * 1.4 + 2, its


def _solver(env, op, precision):
    s = F.iterative_linear_solver_t().create(env)
    s.set_type_from_string("CG")
    s.set_parameters_from_pl({"ILS_RTOL": 1e-10, "ILS_MAX_NUM_ITERATIONS": 20000, "FEMHIP_PRECISION": precision})
    s.set_operators(op, F.jacobi_preconditioner_t(op))
    return s


@pytest.mark.parametrize("name", ["2d_q1_64_sin", "3d_q1_cellk"])
def test_recompute_writes_solver_copies(env, name):
    """Once solvers exist, the owner-computes assembly of a scalar Q1 operator
    also writes their symmetric-diagonal FP64 copy (the FP32 copy is derived
    from it).  A recompute with new coefficients must then solve exactly as a
    fresh operator, whose copies are filled from the CSR."""
    print(f"[recompute] {name}", flush=True)
    osp, oph, fes, op, di = _build(env, name)
    op.compute()
    solvers = {p: _solver(env, op, p) for p in ("FP64", "FP32_REFINED")}
    for s in solvers.values():
        s.solve_resident()
    if di.k == "cell":
        di2 = F.poisson_discrete_integration_t(k="cell", k_cell=di.k_cell * np.linspace(0.5, 2.0, di.k_cell.size),
                                               f=di.f, f0=di.f0)
    else:
        di2 = F.poisson_discrete_integration_t(k="const", k0=3.5, f=di.f, f0=di.f0)
    phys = di2.physics()
    F._native.check(op._lib.femhip_operator_set_physics(op.h, F._native.C.byref(phys)))
    op.compute()
    op_ref = F.fe_affine_operator_t()
    op_ref.create("CSR", [True], fes, di2)
    op_ref.compute()
    assert np.array_equal(op.get_matrix().val, op_ref.get_matrix().val)
    for p, s in solvers.items():
        r = s.solve_resident()
        s_ref = _solver(env, op_ref, p)
        r_ref = s_ref.solve_resident()
        assert r["num_iterations"] == r_ref["num_iterations"], p
        assert np.array_equal(s.get_solution(), s_ref.get_solution()), p


def test_error_paths(env):
    osp, oph, fes, op, di = _build(env, "2d_q1_64_sin")
    op.compute()
    solver = F.iterative_linear_solver_t().create(env)
    solver.set_type_from_string("CG")
    solver.set_parameters_from_pl({"ILS_RTOL": 1e-14, "ILS_MAX_NUM_ITERATIONS": 3})
    solver.set_operators(op, F.jacobi_preconditioner_t(op))
    solver.solve()
    assert not solver.did_converge() and solver.get_num_iterations() == 3
    assert solver.result["status"] == F._native.ERR_NOT_CONVERGED
    # a zero coefficient gives a zero diagonal: Jacobi setup must refuse it
    tria = F.serial_triangulation_t().create(env, {"STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": 2,
                                                   "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": [4, 4]})
    fes0 = F.serial_fe_space_t().create(tria)
    op0 = F.fe_affine_operator_t()
    op0.create("CSR", [True], fes0, F.poisson_discrete_integration_t(k="const", k0=0.0))
    op0.compute()
    s0 = F.iterative_linear_solver_t().create(env)
    s0.set_type_from_string("CG")
    s0.set_operators(op0, F.jacobi_preconditioner_t(op0))
    with pytest.raises(F.FemhipError) as e:
        s0.solve()
    assert e.value.code == F._native.ERR_ZERO_DIAGONAL
    with pytest.raises(F.FemhipError):
        F.iterative_linear_solver_t().create(env).set_type_from_string("GMRES")


def test_drop_in_space_from_reference_dofs(env):
    """The caller's own lst_dofs_gids (FEMPAR's fe_space) drives the same assembly."""
    osp, oph, fes, op, di = _build(env, "3d_q1_aniso")
    tria = fes.triangulation
    fes2 = F.serial_fe_space_t().create(tria, lst_dofs_gids=osp.dofs(), num_free_dofs=osp.n_free,
                                        num_fixed_dofs=osp.n_fixed)
    op2 = F.fe_affine_operator_t()
    op2.create("CSR", [True], fes2, di)
    op.compute()
    op2.compute()
    A, A2 = op.get_matrix(), op2.get_matrix()
    assert np.array_equal(A.irp, A2.irp) and np.array_equal(A.ja, A2.ja) and np.array_equal(A.val, A2.val)


@pytest.mark.parametrize("name", ["3d_q1_20_sin3", "3d_q1_cellk"])
def test_mfma_assembly_matches(env, name, monkeypatch):
    """The matrix-core variant of the 3-D Q1 owner-computes assembly
    (FEMHIP_ASSEMBLY_MFMA=1) gives the oracle's values within the 1e-12 gate,
    the vector kernel's within rounding, and a bitwise symmetric matrix."""
    import scipy.sparse as sps
    print(f"[mfma] {name}", flush=True)
    osp, oph, fes, op, di = _build(env, name, symmetric_storage=False)
    g = osp.fixed_values(oph)
    op.set_fixed_dof_values(g if osp.n_fixed else None)
    op.compute()
    A_vec, b_vec = op.get_matrix(), op.get_translation()
    monkeypatch.setenv("FEMHIP_ASSEMBLY_MFMA", "1")
    op.compute()
    A, b = op.get_matrix(), op.get_translation()
    monkeypatch.delenv("FEMHIP_ASSEMBLY_MFMA")
    irp, ja, val, rhs = osp.assemble(oph, False, fixed_values=g)
    assert np.array_equal(A.irp, irp) and np.array_equal(A.ja, ja)
    scale, bscale = np.max(np.abs(val)), max(np.max(np.abs(rhs)), 1e-300)
    assert np.max(np.abs(A.val - val)) <= VAL_RTOL * scale
    assert np.max(np.abs(b - rhs)) <= VAL_RTOL * bscale
    assert np.max(np.abs(A.val - A_vec.val)) <= 1e-14 * scale
    assert np.max(np.abs(b - b_vec)) <= 1e-14 * bscale
    n = len(irp) - 1
    M = sps.csr_matrix((A.val, A.ja - 1, A.irp - 1), shape=(n, n))
    assert (M != M.T).nnz == 0          # bitwise symmetric


@pytest.mark.parametrize("name", ["2d_q1_64_sin", "3d_q1_cellk", "3d_q1_linear_g"])
def test_owner_assembly_for_q1(env, name, monkeypatch):
    """Scalar Q1 through the general one-wavefront-per-row owner-computes kernel
    (the vertex maps switched off): values and RHS within the 1e-12 gate, the
    matrix bitwise symmetric (every entry summed over the same cells in the
    same order from both of its rows)."""
    import scipy.sparse as sps
    monkeypatch.setenv("FEMHIP_NO_GATHER", "1")
    osp, oph, fes, op, di = _build(env, name, symmetric_storage=False)
    g = osp.fixed_values(oph)
    op.set_fixed_dof_values(fes.interpolate_dirichlet_values(di) if osp.n_fixed else None)
    op.compute()
    monkeypatch.delenv("FEMHIP_NO_GATHER")
    irp, ja, val, rhs = osp.assemble(oph, False, fixed_values=g)
    A = op.get_matrix()
    assert np.array_equal(A.irp, irp) and np.array_equal(A.ja, ja)
    assert np.max(np.abs(A.val - val)) <= VAL_RTOL * np.max(np.abs(val))
    b = op.get_translation()
    assert np.max(np.abs(b - rhs)) <= VAL_RTOL * max(np.max(np.abs(rhs)), 1e-300)
    n = len(irp) - 1
    M = sps.csr_matrix((A.val, A.ja - 1, A.irp - 1), shape=(n, n))
    assert (M != M.T).nnz == 0


def test_q2_assembly_symmetric_and_repeatable(env):
    """3-D Q2: element matrices on the matrix cores, gathered by the rows'
    owners: the matrix is bitwise symmetric and a second compute gives the same
    bits (values already checked against the oracle at 1e-12 above)."""
    import scipy.sparse as sps
    osp, oph, fes, op, di = _build(env, "3d_q2_8_sin3", symmetric_storage=False)
    op.compute()
    A1, b1 = op.get_matrix(), op.get_translation()
    op.compute()
    A2, b2 = op.get_matrix(), op.get_translation()
    assert np.array_equal(A1.val, A2.val) and np.array_equal(b1, b2)
    n = len(A1.irp) - 1
    M = sps.csr_matrix((A1.val, A1.ja - 1, A1.irp - 1), shape=(n, n))
    assert (M != M.T).nnz == 0


@pytest.mark.parametrize("name", ["3d_q2_8_sin3", "2d_q2_12", "2d_q2_linear_g", "3d_q2_x0_7x5x6", "elasticity_8"])
def test_q2_class_layout_solves(env, name, monkeypatch):
    """Q2 and vector Q1: the solver's class-blocked diagonal layout (rows
    renumbered class by class: node parity and component; identity rows in the
    padding) gives the oracle's FP64 solution within
    1e-10 in FP64 and within 1e-8 in FP32-refined, with the iteration count of
    the first-touch ELL layout (FEMHIP_CLASS_DIA=0) within 2."""
    print(f"[q2 layout] {name}", flush=True)
    sols, its = {}, {}
    for layout in ("1", "0"):
        monkeypatch.setenv("FEMHIP_CLASS_DIA", layout)
        osp, oph, fes, op, di = _build(env, name)
        g = osp.fixed_values(oph)
        op.set_fixed_dof_values(g if osp.n_fixed else None)
        op.compute()
        for prec in ("FP64", "FP32_REFINED"):
            solver = F.iterative_linear_solver_t().create(env)
            solver.set_type_from_string("CG")
            solver.set_parameters_from_pl({"ILS_RTOL": 1e-10, "ILS_MAX_NUM_ITERATIONS": 20000,
                                           "FEMHIP_PRECISION": prec})
            solver.set_operators(op.get_tangent(), F.jacobi_preconditioner_t(op))
            sols[layout, prec] = solver.solve(op.get_translation())
            its[layout, prec] = solver.result["num_iterations"]
            assert solver.result["did_converge"], (layout, prec)
    monkeypatch.delenv("FEMHIP_CLASS_DIA")
    irp, ja, val, rhs = osp.assemble(oph, True, fixed_values=g)
    x_ref, info = O.cg(irp, ja, val, rhs, True, precond="jacobi", rtol=1e-12, maxit=20000)
    for layout in ("1", "0"):
        assert np.linalg.norm(sols[layout, "FP64"] - x_ref) <= SOL_RTOL * np.linalg.norm(x_ref), layout
        assert np.linalg.norm(sols[layout, "FP32_REFINED"] - x_ref) <= 1e-8 * np.linalg.norm(x_ref), layout
    assert abs(its["1", "FP64"] - its["0", "FP64"]) <= 2
