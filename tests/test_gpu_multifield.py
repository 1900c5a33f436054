"""Multi-field block operator on the GPU against the oracle, through the C-ABI.

Gates: each field's DOF list in the block numbering identical to the
oracle's; every block's CSR pattern identical (full and symmetric storage of
the diagonal blocks); values and RHS within 1e-12 relative; the block
operator's product within 1e-12; the block CG solution within 1e-10 of the
oracle's CG on the merged matrix, iteration counts within 2.
"""
import numpy as np
import pytest

import fempar_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BC = {"boundary": O.BC_ALL, "x0": O.BC_X0, "none": O.BC_NONE}

CASES = {
    # name: (dim, n, [(order, bc)], field_blocks, coupling, kappa, sigma, f)
    "2d_q2q1_one_block": (2, (9, 7), [(2, "boundary"), (1, "boundary")], [1, 1], None, (1.0, 2.0), 0.75, (1.0, 0.5)),
    "2d_q2q1_two_blocks": (2, (9, 7), [(2, "boundary"), (1, "x0")], [1, 2], None, (1.0, 2.0), 0.75, (1.0, 0.5)),
    "2d_q1q2_uncoupled": (2, (6, 8), [(1, "boundary"), (2, "boundary")], [1, 2], [[1, 0], [0, 1]],
                          (0.5, 1.5), 0.0, (1.0, -1.0)),
    "3d_q1q2_two_blocks": (3, (5, 4, 3), [(1, "boundary"), (2, "boundary")], [2, 1], None, (1.0, 1.0), 3.0,
                           (1.0, 1.0)),
    "3d_three_fields": (3, (4, 4, 4), [(1, "boundary"), (1, "x0"), (2, "boundary")], [1, 2, 1], None,
                        (1.0, 0.5, 2.0), 1.0, (1.0, 0.0, -1.0)),
}


def _build(env, name, symmetric_storage=False):
    dim, n, fields, blocks, coupling, kappa, sigma, f = CASES[name]
    nf = len(fields)
    coupling = np.ones((nf, nf), dtype=bool) if coupling is None else np.asarray(coupling, dtype=bool)
    lim = [0, 1, 0, 1, 0, 1] if dim == 3 else [0, 1, 0, 1, 0, 0]
    probs = [O.make_problem(dim, n, o, bc=BC[bc], lim=lim) for o, bc in fields]
    mf = O.MultiFieldSpace(probs, blocks)
    oph = O.make_mf_physics(kappa, sigma, f)
    tria = F.serial_triangulation_t().create(env, {
        "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": dim, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": list(n),
        "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": lim})
    spaces = [F.serial_fe_space_t().create(tria, order=o, dirichlet=bc) for o, bc in fields]
    layout = F.block_layout_t().create(nf, blocks, coupling)
    op = F.block_fe_affine_operator_t()
    op.create("CSR", [symmetric_storage], spaces, layout,
              F.coupled_diffusion_discrete_integration_t(kappa, sigma, f))
    return mf, oph, coupling, spaces, layout, op


@pytest.mark.parametrize("sym", [False, True])
@pytest.mark.parametrize("name", list(CASES))
def test_block_numbering_pattern_values_rhs(env, name, sym):
    print(f"[multifield] {name} sym={sym}", flush=True)
    mf, oph, coupling, spaces, layout, op = _build(env, name, sym)
    assert op.num_blocks == mf.num_blocks and op.block_sizes == mf.block_sizes
    for fld in range(len(spaces)):
        assert np.array_equal(op.get_lst_dofs_gids(fld + 1), mf.dofs(fld)), fld
    op.compute()
    for ib in range(1, mf.num_blocks + 1):
        for jb in range(1, mf.num_blocks + 1):
            irp, ja, val = mf.assemble_block(oph, coupling, ib, jb, symmetric_storage=sym)
            B = op.get_block(ib, jb)
            assert np.array_equal(B.irp, irp) and np.array_equal(B.ja, ja), (ib, jb)
            if len(val):
                assert np.max(np.abs(B.val - val)) <= 1e-12 * np.max(np.abs(val)), (ib, jb)
    rhs = np.concatenate([mf.assemble_rhs(oph, ib) for ib in range(1, mf.num_blocks + 1)])
    b = op.get_translation()
    assert np.max(np.abs(b - rhs)) <= 1e-12 * max(np.max(np.abs(rhs)), 1e-300)


@pytest.mark.parametrize("prec", ["jacobi", "identity"])
@pytest.mark.parametrize("name", list(CASES))
def test_block_apply_and_cg_match_oracle(env, name, prec):
    print(f"[multifield-cg] {name} {prec}", flush=True)
    mf, oph, coupling, spaces, layout, op = _build(env, name)
    op.compute()
    irp, ja, val, rhs = mf.merged(oph, coupling)
    x = np.random.default_rng(5).standard_normal(len(rhs))
    y_ref = O.spmv(irp, ja, val, x, False)
    assert np.max(np.abs(op.apply(x) - y_ref)) <= 1e-12 * np.max(np.abs(y_ref))
    x_ref, info = O.cg(irp, ja, val, rhs, False, precond=prec, rtol=1e-10, maxit=20000)
    assert info["converged"]
    solver = F.iterative_linear_solver_t().create(env)
    solver.set_type_from_string("CG")
    solver.set_parameters_from_pl({"ILS_RTOL": 1e-10, "ILS_MAX_NUM_ITERATIONS": 20000})
    M = F.jacobi_preconditioner_t(op) if prec == "jacobi" else F.identity_operator_t(op)
    solver.set_operators(op.get_tangent(), M)
    xs = solver.solve()
    r = solver.result
    assert r["did_converge"]
    assert abs(r["num_iterations"] - info["iterations"]) <= 2, (r["num_iterations"], info["iterations"])
    assert np.linalg.norm(xs - x_ref) <= 1e-10 * np.linalg.norm(x_ref)
    assert r["true_residual"] <= 1.01e-10 * np.linalg.norm(
        rhs - O.spmv(irp, ja, val, np.zeros_like(rhs), False))


def test_block_operator_refuses_bad_layouts(env):
    tria = F.serial_triangulation_t().create(env, {
        "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": 2, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": [4, 4],
        "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": [0, 1, 0, 1, 0, 0]})
    other = F.serial_triangulation_t().create(env, {
        "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": 2, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": [5, 4],
        "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": [0, 1, 0, 1, 0, 0]})
    a = F.serial_fe_space_t().create(tria, order=1)
    b = F.serial_fe_space_t().create(other, order=1)
    vec = F.serial_fe_space_t().create(tria, order=1, num_components=2, dirichlet="x0")
    di = F.coupled_diffusion_discrete_integration_t((1.0, 1.0), 1.0, (1.0, 1.0))
    for spaces, layout in (([a, b], F.block_layout_t().create(2, [1, 2])),        # different meshes
                           ([a, vec], F.block_layout_t().create(2, [1, 2])),      # vector field
                           ([a, a], F.block_layout_t().create(2, [1, 2], [[0, 1], [1, 1]]))):   # no self coupling
        with pytest.raises(F.FemhipError) as e:
            F.block_fe_affine_operator_t().create("CSR", [False], spaces, layout, di)
        assert e.value.code == 1
