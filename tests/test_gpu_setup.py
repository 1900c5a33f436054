"""Device-side set-up: first-touch DOF numbering, CSR pattern and Dirichlet values
computed on the GPU (through the C-ABI).

Gates:
  * cell DOF lists, n_free and n_fixed identical to the oracle's first-touch
    numbering (sbm_serial_fe_space.i90:525-715, 2623-2922) on meshes of up to
    ~110k cells, 2-D and 3-D, Q1/Q2, vector Q1, all Dirichlet kinds;
  * the CSR pattern, full and upper (symmetric storage), identical to the
    oracle's COO append + sort_and_compress (base_sparse_matrix.f90:3499-3510);
  * Dirichlet values of the linear g within 1e-14 of the oracle's (different
    but equivalent node-position arithmetic);
  * at 128^3 (2,048,383 free DOFs, too large for the oracle in a test): every
    vertex carries one id in all its cells, ids appear for the first time in
    increasing order along the cell sweep (the definition of first touch), and
    each row's columns are exactly its free neighbour vertices;
  * a caller's numbering that does not conform to the mesh is refused.
"""
import numpy as np
import pytest

import fempar_amd as F
from fempar_amd import _native as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BC = {"boundary": O.BC_ALL, "x0": O.BC_X0, "none": O.BC_NONE}
CASES = [
    (3, (48, 48, 48), 1, 1, "boundary"),
    (3, (37, 23, 11), 1, 1, "boundary"),
    (3, (16, 16, 16), 2, 1, "x0"),
    (3, (14, 14, 14), 1, 3, "x0"),
    (2, (150, 120), 2, 1, "none"),
    (2, (64, 9), 1, 2, "x0"),
    (3, (2, 1, 3), 1, 1, "none"),
    (3, (1, 1, 1), 2, 1, "boundary"),
    (2, (1, 6), 1, 1, "boundary"),          # every DOF fixed
]


def _space(env, dim, n, order, ncomp, bc):
    lim = [0, 1, 0, 1, 0, 1] if dim == 3 else [0, 1, 0, 1, 0, 0]
    tria = F.serial_triangulation_t().create(env, {
        "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": dim, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": list(n),
        "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": lim})
    return F.serial_fe_space_t().create(tria, order=order, num_components=ncomp, dirichlet=bc)


def _physics(ncomp):
    if ncomp > 1:
        return (O.make_physics(kind=O.PHYS_ELASTICITY, E=1.0, nu=0.3, body=(0, 0, -1)),
                F.linear_elasticity_discrete_integration_t(E=1.0, nu=0.3, body=(0, 0, -1)))
    return O.make_physics(), F.poisson_discrete_integration_t(k="const", f="const", f0=1.0, g="linear")


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d_{'x'.join(map(str, c[1]))}_q{c[2]}_c{c[3]}_{c[4]}")
def test_device_numbering_and_pattern(env, case):
    dim, n, order, ncomp, bc = case
    print(f"[setup] {case}", flush=True)
    fes = _space(env, dim, n, order, ncomp, bc)
    osp = O.Space(O.make_problem(dim, n, order, ncomp, BC[bc]))
    assert (fes.num_free_dofs, fes.num_fixed_dofs) == (osp.n_free, osp.n_fixed)
    assert np.array_equal(fes.get_lst_dofs_gids(), osp.dofs())
    oph, di = _physics(ncomp)
    if ncomp == 1 and osp.n_fixed > 0:
        g = fes.interpolate_dirichlet_values(di)
        og = osp.fixed_values(O.make_physics(g_kind=O.G_LINEAR))
        assert np.max(np.abs(g - og)) <= 1e-14 * max(1.0, np.max(np.abs(og)))
    if osp.n_free == 0 or (ncomp > 1 and dim == 2):   # the elasticity plugin is 3-D only
        return
    for sym in (False, True):
        irp, ja, _, _ = osp.assemble(oph, sym)
        op = F.fe_affine_operator_t()
        op.create("CSR", [sym], fes, di)
        a, b = op.get_pattern()
        assert np.array_equal(a, irp) and np.array_equal(b, ja), sym
        op.free()


def test_first_touch_invariants_128(env):
    n = 128
    print("[setup] 128^3 Q1 invariants", flush=True)
    fes = _space(env, 3, (n, n, n), 1, 1, "boundary")
    assert fes.num_free_dofs == (n - 1) ** 3 and fes.num_fixed_dofs == (n + 1) ** 3 - (n - 1) ** 3
    d = fes.get_lst_dofs_gids().astype(np.int64)          # cells x 8, local vertex l = x + 2 y + 4 z bits
    c = np.arange(n ** 3)
    i, j, k = c % n, (c // n) % n, c // (n * n)
    l = np.arange(8)
    vx = i[:, None] + (l & 1)
    vy = j[:, None] + ((l >> 1) & 1)
    vz = k[:, None] + (l >> 2)
    v = (vx + (n + 1) * (vy + (n + 1) * vz)).ravel()
    flat = d.ravel()
    # one id per vertex in all its cells; boundary vertices fixed, interior free
    V = np.zeros((n + 1) ** 3, np.int64)
    V[v] = flat
    assert np.array_equal(V[v], flat)
    vb = ((vx == 0) | (vx == n) | (vy == 0) | (vy == n) | (vz == 0) | (vz == n)).ravel()
    assert np.all((flat < 0) == vb)
    # first touch: along the sweep (cells ascending, local vertices ascending:
    # the polytope order of a Q1 cell's vertices), new ids appear as 1, 2, 3, ...
    for sign, count in ((1, fes.num_free_dofs), (-1, fes.num_fixed_dofs)):
        seq = sign * flat[sign * flat > 0]
        ids, first = np.unique(seq, return_index=True)
        assert np.array_equal(ids, np.arange(1, count + 1)) and np.all(np.diff(first) > 0)
    del d, flat, seq, vb
    # pattern: row of free vertex w = its free neighbours among the 27, ascending
    op = F.fe_affine_operator_t()
    op.create("CSR", [False], fes, F.poisson_discrete_integration_t(k="const", f="const", f0=1.0))
    irp, ja = op.get_pattern()
    Vg = V.reshape(n + 1, n + 1, n + 1)                 # [z, y, x]
    rows = np.full(fes.num_free_dofs, -1, np.int64)
    zz, yy, xx = np.nonzero(Vg > 0)
    rows[Vg[zz, yy, xx] - 1] = (xx + (n + 1) * (yy + (n + 1) * zz))
    assert np.all(rows >= 0)
    del zz, yy, xx
    nb = np.zeros((fes.num_free_dofs, 27), np.int64)
    rz, ry, rx = rows // ((n + 1) ** 2), (rows // (n + 1)) % (n + 1), rows % (n + 1)
    for q in range(27):
        dx, dy, dz = q % 3 - 1, (q // 3) % 3 - 1, q // 9 - 1
        nb[:, q] = Vg[rz + dz, ry + dy, rx + dx]       # interior vertices: neighbours exist
    nb = np.where(nb > 0, nb, np.iinfo(np.int64).max)
    nb.sort(axis=1)
    lens = (nb < np.iinfo(np.int64).max).sum(axis=1)
    assert np.array_equal(np.diff(irp.astype(np.int64)), lens) and irp[0] == 1
    assert np.array_equal(ja.astype(np.int64), nb[nb < np.iinfo(np.int64).max])
    op.free()


def test_nonconforming_numbering_refused(env):
    osp = O.Space(O.make_problem(3, (4, 3, 2), 1, 1, O.BC_ALL))
    tria = F.serial_triangulation_t().create(env, {
        "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS": 3, "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM": [4, 3, 2],
        "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS": [0, 1, 0, 1, 0, 1]})
    d = osp.dofs().copy()
    cell = 1 * 4 * 3 - 1                                 # cell (3, 2, 0): one free vertex
    free = [x for x in range(8) if d[cell, x] > 0]
    other = [x for x in range(8) if d[cell, x] < 0][0]
    d[cell, other] = d[cell, free[0]]                    # the same free id on a second vertex
    with pytest.raises(N.FemhipError) as ei:
        F.serial_fe_space_t().create(tria, lst_dofs_gids=d, num_free_dofs=osp.n_free,
                                     num_fixed_dofs=osp.n_fixed)
    assert ei.value.code == N.ERR_INVALID
