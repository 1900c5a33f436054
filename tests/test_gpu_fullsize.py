"""Parity at the full sizes of the 3-D Poisson configurations and on the
high-contrast random-field workload of the FP32-refined form.

* 128^3 Q1 (2,048,383 free DOFs), k = 1 + 0.5 sin sin sin, manufactured f:
  DOF numbering and symmetric-storage CSR pattern identical to the oracle's,
  values and RHS within 1e-12 relative, the FP64 Jacobi-PCG solution within
  1e-10 of the oracle's FP64 solve (same RES_RES criterion, rtol 1e-10), and
  the FP32-refined solution within 1e-8 of it with a true residual within the
  tolerance.
* 64^3 Q2 (2,048,383 free DOFs): pattern identical, values and RHS within 1e-12.
* Cell-wise k = exp(g), g a Gaussian random field with sigma = 1.5 and
  correlation length 0.1 (contrast above 1e4): FP64 within 1e-10 and
  FP32-refined within 1e-8 of the oracle's FP64 solve, true residuals within
  rtol ||b||.

The oracle takes about a minute for the 128^3 case on one core.
"""
import time

import numpy as np
import pytest

import fempar_amd as F
from bench import grf_k_cell
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _space_and_operator(env, n, order, di, symmetric_storage=True):
    tria = F.serial_triangulation_t().create(env, {
        "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": 3, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": [n, n, n],
        "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": [0, 1, 0, 1, 0, 1]})
    fes = F.serial_fe_space_t().create(tria, order=order)
    op = F.fe_affine_operator_t()
    op.create("CSR", [symmetric_storage], fes, di)
    op.compute()
    return fes, op


def _solver(env, op, precision, rtol):
    s = F.iterative_linear_solver_t().create(env)
    s.set_type_from_string("CG")
    s.set_parameters_from_pl({"ILS_RTOL": rtol, "ILS_MAX_NUM_ITERATIONS": 20000, "FEMHIP_PRECISION": precision})
    s.set_operators(op.get_tangent(), F.jacobi_preconditioner_t(op))
    return s


def _log(msg, t0):
    print(f"[fullsize] {msg} ({time.perf_counter() - t0:.1f} s)", flush=True)


def test_poisson_q1_128_matches_oracle(env):
    t0 = time.perf_counter()
    n = 128
    fes, op = _space_and_operator(env, n, 1, F.poisson_discrete_integration_t(k="sin3", f="manufactured"))
    _log("device space, pattern and assembly done", t0)
    osp = O.Space(O.make_problem(3, (n, n, n), 1))
    assert (fes.num_free_dofs, fes.num_fixed_dofs) == (osp.n_free, osp.n_fixed) == (2048383, 129 ** 3 - 2048383)
    assert np.array_equal(fes.get_lst_dofs_gids(), osp.dofs())
    _log("numbering identical", t0)
    oph = O.make_physics(k_kind=O.K_SIN3, f_kind=O.F_MANUFACTURED)
    irp, ja, val, rhs = osp.assemble(oph, True)
    _log("oracle assembly done", t0)
    A = op.get_matrix()
    assert len(ja) == 28244161
    assert np.array_equal(A.irp, irp) and np.array_equal(A.ja, ja)
    assert np.max(np.abs(A.val - val)) <= 1e-12 * np.max(np.abs(val))
    b = op.get_translation()
    assert np.max(np.abs(b - rhs)) <= 1e-12 * np.max(np.abs(rhs))
    _log("pattern identical, values and RHS within 1e-12", t0)
    del A
    x_ref, info = O.cg(irp, ja, val, rhs, True, precond="jacobi", rtol=1e-10, maxit=5000)
    assert info["converged"]
    _log(f"oracle CG done ({info['iterations']} iterations)", t0)
    del irp, ja, val
    s = _solver(env, op, "FP64", 1e-10)
    x = s.solve(rhs)
    r = s.result
    assert r["did_converge"] and abs(r["num_iterations"] - info["iterations"]) <= 1
    assert np.linalg.norm(x - x_ref) <= 1e-10 * np.linalg.norm(x_ref)
    s32 = _solver(env, op, "FP32_REFINED", 1e-10)
    x32 = s32.solve(rhs)
    r32 = s32.result
    assert r32["did_converge"] and r32["true_residual"] <= 1.0001e-10 * np.linalg.norm(rhs)
    assert np.linalg.norm(x32 - x_ref) <= 1e-8 * np.linalg.norm(x_ref)
    _log(f"FP64 ({r['num_iterations']} it.) and FP32-refined ({r32['num_iterations']} it.) solutions match", t0)
    s.free()
    s32.free()
    op.free()
    fes.free()


def test_poisson_q2_64_pattern_and_values_match_oracle(env):
    t0 = time.perf_counter()
    n = 64
    fes, op = _space_and_operator(env, n, 2, F.poisson_discrete_integration_t(k="sin3", f="manufactured"))
    osp = O.Space(O.make_problem(3, (n, n, n), 2))
    assert fes.num_free_dofs == osp.n_free == 2048383
    irp, ja, val, rhs = osp.assemble(O.make_physics(k_kind=O.K_SIN3, f_kind=O.F_MANUFACTURED), True)
    _log("oracle Q2 assembly done", t0)
    A = op.get_matrix()
    assert len(ja) == 64655955
    assert np.array_equal(A.irp, irp) and np.array_equal(A.ja, ja)
    assert np.max(np.abs(A.val - val)) <= 1e-12 * np.max(np.abs(val))
    b = op.get_translation()
    assert np.max(np.abs(b - rhs)) <= 1e-12 * np.max(np.abs(rhs))
    _log("Q2 pattern identical, values and RHS within 1e-12", t0)
    op.free()
    fes.free()


@pytest.mark.parametrize("n", [32, 40])
def test_random_field_contrast_solves_match_oracle(env, n):
    """The mixed-precision workload at reduced size: k per cell = exp(mean g)
    with the same generator, sigma and correlation length as the 192^3 bench
    configuration, f = 1.  The FP32 inner iterations run on S A S with
    S = diag(A)^(-1/2); the reliable updates must still deliver the FP64
    accuracy on this contrast, also with the FP32 matrix integrated and
    summed in FP32 (FP32_STRICT)."""
    t0 = time.perf_counter()
    kc = grf_k_cell(n)
    assert kc.max() / kc.min() > 1e4
    di = F.poisson_discrete_integration_t(k="cell", k_cell=kc, f="const", f0=1.0)
    fes, op = _space_and_operator(env, n, 1, di)
    osp = O.Space(O.make_problem(3, (n, n, n), 1))
    irp, ja, val, rhs = osp.assemble(O.make_physics(k_kind=O.K_CELL, k_cell=kc, f_kind=O.F_CONST, f0=1.0), True)
    A = op.get_matrix()
    assert np.array_equal(A.ja, ja) and np.max(np.abs(A.val - val)) <= 1e-12 * np.max(np.abs(val))
    x_ref, info = O.cg(irp, ja, val, rhs, True, precond="jacobi", rtol=1e-12, maxit=20000)
    assert info["converged"]
    _log(f"{n}^3 random field: contrast {kc.max() / kc.min():.1e}, oracle {info['iterations']} it.", t0)
    bn = np.linalg.norm(rhs)
    for prec, tol in (("FP64", 1e-10), ("FP32_REFINED", 1e-8), ("FP32_STRICT", 1e-8)):
        s = _solver(env, op, prec, 1e-10)
        x = s.solve(rhs)
        r = s.result
        assert r["did_converge"], prec
        assert r["true_residual"] <= 1.0001e-10 * bn, (prec, r["true_residual"] / bn)
        err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
        assert err <= tol, (prec, err)
        _log(f"{prec}: {r['num_iterations']} it., {r['outer_iterations']} updates, rel. error {err:.2e}", t0)
        s.free()
    op.free()
    fes.free()
