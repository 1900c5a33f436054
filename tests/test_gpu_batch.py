"""Batched form: every problem of a mixed batch assembled in one launch and
solved in one persistent launch, checked against the oracle per problem.

Gates: CSR pattern bit-exact and values/RHS within 1e-12 relative of the
oracle's assembly of the same mesh (2-D/3-D Q1, constant k, manufactured
source); FP64 true residual <= rtol ||b|| (checked on the host with the
oracle's matrix); solution within 1e-5 relative of the oracle's FP64 CG
solution to 1e-12 (the FP32-refined solve stops at rtol = 1e-8)."""
import numpy as np
import pytest

import fempar_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PROBLEMS = [(2, 4, 1.0, 1.0), (2, 7, 0.1, 1.0), (3, 3, 2.5, 1.0), (2, 16, 9.0, 1.0), (3, 5, 0.7, 2.0),
            (2, 33, 3.3, 1.0), (3, 8, 1.2, 1.0), (2, 2, 1.0, 1.0), (3, 11, 4.0, 0.5), (2, 64, 0.5, 1.0),
            (3, 16, 7.5, 1.0), (2, 100, 1.0, 1.0), (3, 2, 3.0, 1.0)]   # n = 2: one row, aliased offsets


def _oracle(dim, n, k, f0):
    sp = O.Space(O.make_problem(dim, [n] * dim, 1))
    ph = O.make_physics(k_kind=O.K_CONST, k0=k, f_kind=O.F_CELL_SIN, f0=f0)
    return sp.assemble(ph, symmetric_storage=False)


@pytest.fixture(scope="module")
def batch(env):
    b = F.batch_t(env, PROBLEMS, rtol=1e-8)
    b.assemble()
    yield b
    b.free()


def test_batch_matrices_match_oracle(batch):
    for i, (dim, n, k, f0) in enumerate(PROBLEMS):
        irp, ja, val, rhs = _oracle(dim, n, k, f0)
        birp, bja, bval, brhs = batch.get_problem_matrix(i)
        assert np.array_equal(birp, irp) and np.array_equal(bja, ja), (dim, n)
        if len(val):
            assert np.max(np.abs(bval - val)) <= 1e-12 * np.max(np.abs(val)), (dim, n)
            assert np.max(np.abs(brhs - rhs)) <= 1e-12 * np.max(np.abs(rhs)), (dim, n)


def test_batch_solutions(batch):
    for rep in range(2):        # the hipGraph is replayed: same answers
        print(f"[batch] solve replay {rep} of {len(PROBLEMS)} problems", flush=True)
        res = batch.solve()
        print(f"[batch] replay {rep}: {res['ms']:.3f} ms, iterations {res['num_iterations'].tolist()}", flush=True)
        assert np.all(res["did_converge"] == 1)
        assert np.all(res["latency_us"] > 0) and res["ms"] > 0
        for i, (dim, n, k, f0) in enumerate(PROBLEMS):
            irp, ja, val, rhs = _oracle(dim, n, k, f0)
            x = batch.get_solution(i)
            r = rhs - O.spmv(irp, ja, val, x, False)
            assert np.linalg.norm(r) <= 1.0001e-8 * np.linalg.norm(rhs), (dim, n)
            assert abs(res["true_residual"][i] - np.linalg.norm(r)) <= 1e-6 * np.linalg.norm(rhs) + 1e-300
            x_ref, info = O.cg(irp, ja, val, rhs, False, precond="jacobi", rtol=1e-12, maxit=20000)
            assert np.linalg.norm(x - x_ref) <= 1e-5 * np.linalg.norm(x_ref), (dim, n)
        if rep == 0:
            first = [batch.get_solution(i) for i in range(len(PROBLEMS))]
        else:
            for i in range(len(PROBLEMS)):
                assert np.array_equal(first[i], batch.get_solution(i))


def test_batch_autotune(batch):
    print("[batch] autotune", flush=True)
    t = batch.autotune()
    assert t in (256, 512, 1024)
    res = batch.solve()
    assert np.all(res["did_converge"] == 1)
