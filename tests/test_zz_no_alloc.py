"""femhip_operator_apply_noalloc against femhip_operator_apply, through the C-ABI.

The no-allocation entry point stages x and y in a fixed device array shared by
every call on the GPU; two calls in a row (the second on a smaller operator,
so stale values of the first sit beyond its rows) must each give the allocating
entry point's result bit for bit, and an operator beyond the array's 2,097,152
rows is refused with FEMHIP_ERR_INVALID.
"""
import numpy as np
import pytest

import fempar_amd as F
from fempar_amd import _native as N

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 21


def _operator(env, dim, n, order=1, k="sin3"):
    lim = [0, 1, 0, 1, 0, 1] if dim == 3 else [0, 1, 0, 1, 0, 0]
    tria = F.serial_triangulation_t().create(env, {
        "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": dim, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": list(n),
        "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": lim})
    fes = F.serial_fe_space_t().create(tria, order=order)
    op = F.fe_affine_operator_t()
    op.create("CSR", [False], fes, F.poisson_discrete_integration_t(k=k, f="manufactured"))
    op.compute()
    return fes, op


def _apply_noalloc(op, x):
    lib = N.load()
    y = np.full(op.n, np.nan)
    N.check(lib.femhip_operator_apply_noalloc(op.h, N.ptr(x), N.ptr(y)))
    return y


@pytest.mark.parametrize("seed", [11])
def test_noalloc_apply_matches_allocating_apply_twice(env, seed):
    rng = np.random.default_rng(seed)
    print("[noalloc] 3-D Q1 48^3 then 2-D Q2 20^2", flush=True)
    fes1, op1 = _operator(env, 3, (48, 48, 48))
    fes2, op2 = _operator(env, 2, (20, 20), order=2)
    assert op2.n < op1.n <= MAX_ROWS
    for op in (op1, op2):                  # larger first, then smaller
        x = rng.standard_normal(op.n)
        y_ref = op.apply(x)
        y = _apply_noalloc(op, x)
        assert np.array_equal(y, y_ref)
    # and back to the larger one after the smaller call
    x = rng.standard_normal(op1.n)
    assert np.array_equal(_apply_noalloc(op1, x), op1.apply(x))


def test_noalloc_apply_refuses_operators_beyond_its_maximum(env):
    print("[noalloc] 3-D Q1 130^3 (2,146,689 rows) refused", flush=True)
    fes, op = _operator(env, 3, (130, 130, 130), k="const")
    assert op.n > MAX_ROWS
    lib = N.load()
    x = np.zeros(op.n)
    y = np.zeros(op.n)
    rc = lib.femhip_operator_apply_noalloc(op.h, N.ptr(x), N.ptr(y))
    assert rc == N.ERR_INVALID
    assert b"maximum" in lib.femhip_last_error()
    # the allocating entry point still serves it
    x[:] = 1.0
    assert np.all(np.isfinite(op.apply(x)))


def test_noalloc_apply_rejects_null_arguments():
    lib = N.load()
    assert lib.femhip_operator_apply_noalloc(None, None, None) == N.ERR_INVALID
    assert b"null" in lib.femhip_last_error()
