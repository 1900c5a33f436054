"""Host-side mirror of FEMPAR's interfaces for the one-subdomain hot path.

Names, argument meaning and error behaviour follow the Fortran API
(Sources/Lib/...), so driver code reads like the reference's
test_poisson_driver.f90:

    tria = serial_triangulation_t(); tria.create(env, plist)
    fe_space = serial_fe_space_t(); fe_space.create(tria, order=1)
    op = fe_affine_operator_t()
    op.create('CSR', [True], fe_space, poisson_discrete_integration_t(...))
    op.compute()
    solver = iterative_linear_solver_t(); solver.create(env)
    solver.set_type_from_string('CG'); solver.set_parameters_from_pl(plist)
    solver.set_operators(op.get_tangent(), jacobi_preconditioner_t(op))
    x = solver.solve(op.get_translation())

Everything below is plumbing over the C-ABI (include/femhip.h); the work runs
in libfemhip's HIP kernels.  A failing call raises FemhipError, the analogue of
the reference's check/mcheck abort.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from ._native import FemhipError

# ParameterList keys (uniform_hex_mesh_generator_parameters.f90:31-35,
# iterative_solver_parameters.f90:76-83)
STRUCT_HEX_MESH_GENERATOR_NUM_DIMS = "STRUCT_HEX_MESH_GENERATOR_NUM_DIMS"
STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM = "STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM"
STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS = "STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS"
ILS_TYPE = "ILS_TYPE"
ILS_RTOL = "ILS_RTOL"
ILS_ATOL = "ILS_ATOL"
ILS_STOPPING_CRITERIUM = "ILS_STOPPING_CRITERIUM"
ILS_MAX_NUM_ITERATIONS = "ILS_MAX_NUM_ITERATIONS"
# keys of this implementation's refined mode (no reference counterpart)
ILS_PRECISION = "FEMHIP_PRECISION"          # "FP64" | "FP32_REFINED" | "FP32_STRICT"
ILS_INNER_RTOL = "FEMHIP_INNER_RTOL"
ILS_MAX_OUTER = "FEMHIP_MAX_OUTER"

# defaults of iterative_solver_parameters.f90:166-178
DEFAULT_RTOL = 1.0e-6
DEFAULT_ATOL = 0.0
DEFAULT_MAX_NUM_ITERATIONS = 1000
DEFAULT_CG_STOPPING_CRITERIA = "RES_RES"

cg_name = "CG"


class environment_t:
    """One GPU per process: owns the femhip context (device + stream)."""

    def __init__(self, device: int = 0):
        self._lib = N.load()
        h = C.c_void_p()
        N.check(self._lib.femhip_ctx_create(device, C.byref(h)))
        self.h = h

    def free(self):
        if getattr(self, "h", None):
            self._lib.femhip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class serial_triangulation_t:
    """Struct-hex triangulation (sbm_serial_triangulation.i90:61-118); only the
    generator parameters are kept, the mesh is implicit.  vertex_coordinates
    (num_vertices x num_dims, vertex (i, j, k) at row i + (n0+1)(j + (n1+1) k))
    moves the vertices: the topology stays the structured one and every cell
    is integrated with its own map."""

    def create(self, environment: environment_t, parameter_list: dict, vertex_coordinates=None):
        self.environment = environment
        self.num_dims = int(parameter_list.get(STRUCT_HEX_MESH_GENERATOR_NUM_DIMS, 3))
        n = list(parameter_list.get(STRUCT_HEX_MESH_GENERATOR_NUM_CELLS_X_DIM, [10, 10, 10]))
        self.num_cells_x_dim = (n + [1, 1, 1])[:3]
        lim = list(parameter_list.get(STRUCT_HEX_MESH_GENERATOR_DOMAIN_LIMITS, [0.0, 1.0, 0.0, 1.0, 0.0, 1.0]))
        self.domain_limits = (lim + [0.0] * 6)[:6]
        self.vertex_coordinates = None
        if vertex_coordinates is not None:
            vc = np.ascontiguousarray(vertex_coordinates, dtype=np.float64).reshape(-1, self.num_dims)
            if vc.shape[0] != self.get_num_vertices():
                raise FemhipError(N.ERR_INVALID, "vertex_coordinates must have one row per vertex")
            self.vertex_coordinates = vc
        return self

    def get_num_dims(self):
        return self.num_dims

    def get_num_vertices(self):
        m = 1
        for k in range(self.num_dims):
            m *= self.num_cells_x_dim[k] + 1
        return m

    def get_uniform_vertex_coordinates(self):
        """The generator's vertex coordinates (uniform_hex_mesh_generator.f90:608-612,
        633-635), num_vertices x num_dims, x fastest."""
        axes = []
        for k in range(self.num_dims):
            nk = self.num_cells_x_dim[k]
            a, b = self.domain_limits[2 * k], self.domain_limits[2 * k + 1]
            axes.append((b - a) * (np.arange(nk + 1) / nk) + a)
        grids = np.meshgrid(*axes, indexing="ij")
        return np.stack([g.transpose().ravel() for g in grids], axis=1)


class serial_fe_space_t:
    """serial_fe_space_t on a struct-hex triangulation: strong Dirichlet and
    first-touch DOF numbering (sbm_serial_fe_space.i90:189-244, 525-715, 2623-2922)."""

    def create(self, triangulation: serial_triangulation_t, order: int = 1, num_components: int = 1,
               dirichlet: str = "boundary", lst_dofs_gids=None, num_free_dofs=None, num_fixed_dofs=None):
        self.triangulation = triangulation
        env = triangulation.environment
        self._lib = N.load()
        d = N.MeshDesc()
        d.dim = triangulation.num_dims
        for i in range(3):
            d.n[i] = triangulation.num_cells_x_dim[i]
        for i in range(6):
            d.limits[i] = triangulation.domain_limits[i]
        d.order, d.num_components = order, num_components
        d.dirichlet = {"boundary": N.BC_ALL_BOUNDARY, "x0": N.BC_X0_FACE, "none": N.BC_NONE}[dirichlet]
        self.desc = d
        h = C.c_void_p()
        if lst_dofs_gids is None:
            N.check(self._lib.femhip_space_create_structured(env.h, C.byref(d), C.byref(h)))
        else:
            arr = np.ascontiguousarray(lst_dofs_gids, dtype=np.int32)
            N.check(self._lib.femhip_space_create_from_dofs(env.h, C.byref(d), N.ptr(arr), int(num_free_dofs),
                                                            int(num_fixed_dofs), C.byref(h)))
        self.h = h
        if triangulation.vertex_coordinates is not None:
            N.check(self._lib.femhip_space_set_vertex_coords(h, N.ptr(triangulation.vertex_coordinates)))
        nc, nd, nq, nf, nx = C.c_int64(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        N.check(self._lib.femhip_space_info(h, C.byref(nc), C.byref(nd), C.byref(nq), C.byref(nf), C.byref(nx)))
        self.num_cells, self.num_dofs_x_cell, self.num_quadrature_points = nc.value, nd.value, nq.value
        self.num_free_dofs, self.num_fixed_dofs = nf.value, nx.value
        return self

    def get_num_dofs(self):
        return self.num_free_dofs

    def get_num_fixed_dofs(self):
        return self.num_fixed_dofs

    def get_lst_dofs_gids(self):
        out = np.empty((self.num_cells, self.num_dofs_x_cell), dtype=np.int32)
        N.check(self._lib.femhip_space_get_dofs(self.h, N.ptr(out)))
        return out

    def get_quadrature_points_coordinates(self):
        """fe_cell_iterator_t%get_quadrature_points_coordinates for every cell:
        (num_cells, num_quadrature_points, num_dims)."""
        d = self.triangulation.num_dims
        out = np.empty((self.num_cells, self.num_quadrature_points, d), dtype=np.float64)
        N.check(self._lib.femhip_space_quadrature_points(self.h, N.ptr(out)))
        return out

    def interpolate_dirichlet_values(self, discrete_integration):
        out = np.zeros(max(1, self.num_fixed_dofs))
        p = N.Physics()   # only the Dirichlet function g is read
        p.g_kind = {"zero": N.G_ZERO, "linear": N.G_LINEAR}[getattr(discrete_integration, "g", "zero")]
        N.check(self._lib.femhip_space_interpolate_dirichlet(self.h, C.byref(p), N.ptr(out)))
        return out[: self.num_fixed_dofs]

    def free(self):
        if getattr(self, "h", None):
            self._lib.femhip_space_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class poisson_discrete_integration_t:
    """-div(k grad u) = f (poisson_discrete_integration.f90:62-141 with k(x_q)).
    k and f may also be callables of the quadrature points' coordinates (an
    array (..., num_dims) -> values): like the reference plugin's
    source_term%get_value(quad_coords(qpoint)) (:109-128), they are evaluated
    at every cell's quadrature points and passed as point values."""

    def __init__(self, k="const", k0=1.0, k_cell=None, f="const", f0=0.0, g="zero"):
        self.k, self.k0, self.k_cell, self.f, self.f0, self.g = k, k0, k_cell, f, f0, g

    def physics(self, fe_space=None):
        p = N.Physics()
        p.kind = N.PHYS_POISSON
        self._keep = []
        xq = None
        if callable(self.k) or callable(self.f):
            if fe_space is None:
                raise FemhipError(N.ERR_INVALID, "a callable coefficient needs the FE space's quadrature points")
            xq = fe_space.get_quadrature_points_coordinates()

        def point_values(fn):
            v = np.ascontiguousarray(np.asarray(fn(xq), dtype=np.float64).reshape(-1))
            if v.size != xq.shape[0] * xq.shape[1]:
                raise FemhipError(N.ERR_INVALID, "a callable coefficient must give one value per quadrature point")
            self._keep.append(v)
            return v.ctypes.data_as(C.POINTER(C.c_double))

        if callable(self.k):
            p.k_kind = N.K_QPOINT
            p.k_qp = point_values(self.k)
        else:
            p.k_kind = {"const": N.K_CONST, "sin3": N.K_SIN3, "cell": N.K_CELL}[self.k]
        p.k0 = self.k0
        if self.k_cell is not None:
            kc = np.ascontiguousarray(self.k_cell, dtype=np.float64)
            self._keep.append(kc)
            p.k_cell = kc.ctypes.data_as(C.POINTER(C.c_double))
        if callable(self.f):
            p.f_kind = N.F_QPOINT
            p.f_qp = point_values(self.f)
        else:
            p.f_kind = {"const": N.F_CONST, "manufactured": N.F_MANUFACTURED, "cell_sin": N.F_CELL_SIN}[self.f]
        p.f0 = self.f0
        p.g_kind = {"zero": N.G_ZERO, "linear": N.G_LINEAR}[self.g]
        return p


class linear_elasticity_discrete_integration_t:
    """Irreducible linear elasticity (sbm_irreducible_discrete_integration.i90:64-130)."""

    def __init__(self, E=1.0, nu=0.2, body=(0.0, 0.0, -1.0)):
        self.E, self.nu, self.body = E, nu, body

    def physics(self, fe_space=None):
        p = N.Physics()
        p.kind = N.PHYS_ELASTICITY
        p.E, p.nu = self.E, self.nu
        for i in range(3):
            p.body[i] = self.body[i]
        return p


class csr_matrix_view:
    """Host copy of csr_sparse_matrix_t's irp/ja/val (1-based like the reference)."""

    def __init__(self, irp, ja, val, symmetric_storage):
        self.irp, self.ja, self.val, self.symmetric_storage = irp, ja, val, symmetric_storage

    def get_num_rows(self):
        return len(self.irp) - 1

    def get_nnz(self):
        return len(self.ja)


class fe_affine_operator_t:
    """fe_affine_operator_t with a CSR sparse_assembler_t (fe_operator.f90:174-183)."""

    def create(self, sparse_matrix_storage_format, diagonal_blocks_symmetric_storage, fe_space: serial_fe_space_t,
               discrete_integration, diagonal_blocks_symmetric=None, diagonal_blocks_sign=None):
        if sparse_matrix_storage_format != "CSR":
            raise FemhipError(N.ERR_INVALID, "only CSR storage is supported")
        self.fe_space = fe_space
        self.discrete_integration = discrete_integration
        self.symmetric_storage = bool(diagonal_blocks_symmetric_storage[0])
        self._lib = N.load()
        h = C.c_void_p()
        N.check(self._lib.femhip_operator_create(fe_space.triangulation.environment.h, fe_space.h,
                                                 int(self.symmetric_storage), C.byref(h)))
        self.h = h
        n, nnz, nnzf = C.c_int32(), C.c_int64(), C.c_int64()
        N.check(self._lib.femhip_operator_info(h, C.byref(n), C.byref(nnz), C.byref(nnzf)))
        self.n, self.nnz, self.nnz_full = n.value, nnz.value, nnzf.value
        self._phys = discrete_integration.physics(fe_space)
        N.check(self._lib.femhip_operator_set_physics(h, C.byref(self._phys)))
        self.fixed_values = None
        self.last_compute_ms = None
        return self

    def set_fixed_dof_values(self, values):
        self.fixed_values = None if values is None else np.ascontiguousarray(values, dtype=np.float64)

    def compute(self):
        ms = C.c_double()
        N.check(self._lib.femhip_operator_compute(self.h, N.ptr(self.fixed_values), C.byref(ms)))
        self.last_compute_ms = ms.value

    def get_pattern(self):
        irp = np.empty(self.n + 1, dtype=np.int32)
        ja = np.empty(self.nnz, dtype=np.int32)
        N.check(self._lib.femhip_operator_get_pattern(self.h, N.ptr(irp), N.ptr(ja)))
        return irp, ja

    def get_matrix(self) -> csr_matrix_view:
        irp, ja = self.get_pattern()
        val = np.empty(self.nnz)
        N.check(self._lib.femhip_operator_get_values(self.h, N.ptr(val), None))
        return csr_matrix_view(irp, ja, val, self.symmetric_storage)

    def get_translation(self):
        rhs = np.empty(self.n)
        N.check(self._lib.femhip_operator_get_values(self.h, None, N.ptr(rhs)))
        return rhs

    def get_tangent(self):
        return self

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        N.check(self._lib.femhip_operator_apply(self.h, N.ptr(x), N.ptr(y)))
        return y

    def free(self):
        if getattr(self, "h", None):
            self._lib.femhip_operator_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class identity_operator_t:
    """.identity. operator (operator.f90:39-180) used as M."""

    kind = N.PREC_IDENTITY

    def __init__(self, op=None):
        self.op = op


class jacobi_preconditioner_t:
    """Jacobi preconditioner M = diag(A)^-1 (jacobi_preconditioner.f90:136-450).
    Unlike the reference, it also works on a serial FE space."""

    kind = N.PREC_JACOBI

    def __init__(self, op=None):
        self.op = op


class iterative_linear_solver_t:
    """iterative_linear_solver_t restricted to 'CG' (iterative_linear_solver.f90:62-241)."""

    def create(self, environment: environment_t):
        self.environment = environment
        self.type = None
        self.params = {ILS_RTOL: DEFAULT_RTOL, ILS_ATOL: DEFAULT_ATOL,
                       ILS_MAX_NUM_ITERATIONS: DEFAULT_MAX_NUM_ITERATIONS,
                       ILS_STOPPING_CRITERIUM: DEFAULT_CG_STOPPING_CRITERIA, ILS_PRECISION: "FP64",
                       ILS_INNER_RTOL: 0.0, ILS_MAX_OUTER: 0}
        self.A = None
        self.M = None
        self.h = None
        self._lib = N.load()
        return self

    def set_type_from_string(self, name: str):
        if name != cg_name:
            raise FemhipError(N.ERR_INVALID, f"iterative linear solver '{name}' is not available (only CG)")
        self.type = name

    def set_parameters_from_pl(self, parameter_list: dict):
        if ILS_TYPE in parameter_list:
            self.set_type_from_string(parameter_list[ILS_TYPE])
        for k in self.params:
            if k in parameter_list:
                self.params[k] = parameter_list[k]
        if self.params[ILS_STOPPING_CRITERIUM] not in N.STOPPING_CRITERIA:
            raise FemhipError(N.ERR_INVALID, f"CG does not support {self.params[ILS_STOPPING_CRITERIUM]}")
        self._destroy()

    def set_operators(self, A: fe_affine_operator_t, M):
        self.A, self.M = A, M
        self._destroy()

    def _ensure(self):
        if self.type is None:
            raise FemhipError(N.ERR_STATE, "set_type_from_string not called")
        if self.A is None:
            raise FemhipError(N.ERR_STATE, "set_operators not called")
        if self.h is None:
            p = N.CgParams()
            p.rtol = float(self.params[ILS_RTOL])
            p.atol = float(self.params[ILS_ATOL])
            p.max_num_iterations = int(self.params[ILS_MAX_NUM_ITERATIONS])
            p.stopping_criterion = N.STOPPING_CRITERIA.index(self.params[ILS_STOPPING_CRITERIUM])
            p.preconditioner = getattr(self.M, "kind", N.PREC_IDENTITY)
            p.precision = {"FP64": N.PRECISION_FP64, "FP32_REFINED": N.PRECISION_FP32_REFINED,
                           "FP32_STRICT": N.PRECISION_FP32_STRICT}[self.params[ILS_PRECISION]]
            p.inner_rtol = float(self.params[ILS_INNER_RTOL])
            p.max_outer = int(self.params[ILS_MAX_OUTER])
            h = C.c_void_p()
            N.check(self._lib.femhip_solver_create(self.environment.h, self.A.h, C.byref(p), C.byref(h)))
            self.h = h

    def _cg_params(self):
        p = N.CgParams()
        p.rtol = float(self.params[ILS_RTOL])
        p.atol = float(self.params[ILS_ATOL])
        p.max_num_iterations = int(self.params[ILS_MAX_NUM_ITERATIONS])
        p.stopping_criterion = N.STOPPING_CRITERIA.index(self.params[ILS_STOPPING_CRITERIUM])
        p.preconditioner = getattr(self.M, "kind", N.PREC_IDENTITY)
        p.precision = {"FP64": N.PRECISION_FP64, "FP32_REFINED": N.PRECISION_FP32_REFINED,
                       "FP32_STRICT": N.PRECISION_FP32_STRICT}[self.params[ILS_PRECISION]]
        p.inner_rtol = float(self.params[ILS_INNER_RTOL])
        p.max_outer = int(self.params[ILS_MAX_OUTER])
        return p

    def solve(self, b=None, x=None):
        """Solve A x = b from the initial guess x (zero if None); b None uses the
        operator's assembled RHS on the device.  Returns x (host).  A block
        operator is solved over its blocks' rows in block order."""
        if isinstance(self.A, block_fe_affine_operator_t):
            if self.type is None:
                raise FemhipError(N.ERR_STATE, "set_type_from_string not called")
            n = self.A.num_rows
            bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
            xx = np.zeros(n) if x is None else np.array(x, dtype=np.float64)
            res = N.CgResult()
            p = self._cg_params()
            N.check(self._lib.femhip_block_operator_solve(self.A.h, C.byref(p), N.ptr(bb), N.ptr(xx), C.byref(res)))
            self.result = {f: getattr(res, f) for f, _ in N.CgResult._fields_}
            return xx
        self._ensure()
        n = self.A.n
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        xx = np.zeros(n) if x is None else np.array(x, dtype=np.float64)
        res = N.CgResult()
        N.check(self._lib.femhip_solver_solve(self.h, N.ptr(bb), N.ptr(xx), C.byref(res)))
        self.result = {f: getattr(res, f) for f, _ in N.CgResult._fields_}
        return xx

    def solve_resident(self):
        """Solve with the operator's device RHS from x0 = 0, solution kept on the device."""
        self._ensure()
        res = N.CgResult()
        N.check(self._lib.femhip_solver_solve(self.h, None, None, C.byref(res)))
        self.result = {f: getattr(res, f) for f, _ in N.CgResult._fields_}
        return self.result

    def get_solution(self):
        x = np.empty(self.A.n)
        N.check(self._lib.femhip_solver_get_solution(self.h, N.ptr(x)))
        return x

    def get_error_estimate_history_convergence_test(self):
        """error_estimate_history_convergence_test of the last FP64 solve: the
        quantity the stopping criterion compared at each iteration
        (base_iterative_linear_solver.f90:520-525, filled at cg.f90:388-389)."""
        its = C.c_int32()
        N.check(self._lib.femhip_solver_get_convergence_history(self.h, 0, None, C.byref(its)))
        h = np.empty(max(its.value, 1))
        N.check(self._lib.femhip_solver_get_convergence_history(self.h, its.value, N.ptr(h), C.byref(its)))
        return h[: its.value]

    def error_norms(self, exact="sin"):
        """error_norms_scalar_t%compute of the last solution: l2_norm, h1_seminorm,
        h1_norm, linfty_norm against u = prod sin(pi x_d) ("sin") or x + y (+ z) ("linear")."""
        r = N.ErrorNorms()
        N.check(self._lib.femhip_solver_error_norms(self.h, {"sin": N.EXACT_SIN, "linear": N.EXACT_LINEAR}[exact],
                                                    C.byref(r)))
        return {f: getattr(r, f) for f, _ in N.ErrorNorms._fields_}

    def profile_spmv(self, reps=50):
        """Average duration (ms) and algorithmic bytes of one SpMV launch of this solver."""
        self._ensure()
        ms, by = C.c_double(), C.c_double()
        N.check(self._lib.femhip_solver_profile_spmv(self.h, int(reps), C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def profile_cg(self):
        """Mean SpMV duration (ms) inside one solve, iterations, algorithmic bytes per SpMV."""
        self._ensure()
        ms, it, by = C.c_double(), C.c_int32(), C.c_double()
        N.check(self._lib.femhip_solver_profile_cg(self.h, C.byref(ms), C.byref(it), C.byref(by)))
        return ms.value, it.value, by.value

    def get_num_iterations(self):
        return self.result["num_iterations"]

    def did_converge(self):
        return bool(self.result["did_converge"])

    def _destroy(self):
        if getattr(self, "h", None):
            self._lib.femhip_solver_destroy(self.h)
            self.h = None

    def free(self):
        self._destroy()

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass


class block_layout_t:
    """block_layout_t (Integration/FEOperator/block_layout.f90:36-264): which
    block each field's DOFs go to (1-based) and which field pairs couple."""

    def create(self, num_fields, field_id_to_block_id=None, field_coupling=None):
        if not 1 <= num_fields <= N.MAX_FIELDS:
            raise FemhipError(N.ERR_INVALID, f"num_fields must be in [1, {N.MAX_FIELDS}]")
        self.num_fields = num_fields
        self.field_id_to_block_id = list(field_id_to_block_id) if field_id_to_block_id is not None \
            else [1] * num_fields
        self.field_coupling = np.ones((num_fields, num_fields), dtype=bool) if field_coupling is None \
            else np.asarray(field_coupling, dtype=bool).reshape(num_fields, num_fields)
        self.num_blocks = max(self.field_id_to_block_id)
        # block_layout_create (block_layout.f90:106-113): blocks couple where any of their fields do
        self.block_coupling = np.zeros((self.num_blocks, self.num_blocks), dtype=bool)
        for i in range(num_fields):
            for j in range(num_fields):
                self.block_coupling[self.field_id_to_block_id[i] - 1, self.field_id_to_block_id[j] - 1] |= \
                    self.field_coupling[i, j]
        return self

    def get_num_fields(self):
        return self.num_fields

    def get_num_blocks(self):
        return self.num_blocks

    def get_block_id(self, field_id):
        return self.field_id_to_block_id[field_id - 1]

    def fields_coupled(self, ifield, jfield):
        return bool(self.field_coupling[ifield - 1, jfield - 1])

    def blocks_coupled(self, iblock, jblock):
        return bool(self.block_coupling[iblock - 1, jblock - 1])


class coupled_diffusion_discrete_integration_t:
    """Multi-field plugin: sum_f kappa_f (grad u_f, grad v_f) + sigma sum_{f<g}
    (u_f - u_g, v_f - v_g) = sum_f (f_f, v_f)."""

    def __init__(self, kappa, sigma=0.0, f=None):
        self.kappa, self.sigma = list(kappa), float(sigma)
        self.f = list(f) if f is not None else [0.0] * len(self.kappa)

    def physics(self):
        p = N.MultiFieldPhysics()
        for i in range(N.MAX_FIELDS):
            p.kappa[i] = self.kappa[i] if i < len(self.kappa) else 0.0
            p.f[i] = self.f[i] if i < len(self.f) else 0.0
        p.sigma = self.sigma
        return p


class block_fe_affine_operator_t:
    """fe_affine_operator_t over several fields with a block_sparse_assembler_t
    (Integration/FEOperator/block_sparse_assembler.f90:47-224): one
    block_sparse_matrix_t block per coupled block pair and one RHS block per block."""

    def create(self, sparse_matrix_storage_format, diagonal_blocks_symmetric_storage, fe_spaces, block_layout,
               discrete_integration):
        if sparse_matrix_storage_format != "CSR":
            raise FemhipError(N.ERR_INVALID, "only CSR storage is supported")
        self._lib = N.load()
        self.fe_spaces, self.block_layout = list(fe_spaces), block_layout
        self.symmetric_storage = bool(diagonal_blocks_symmetric_storage[0])
        nf = block_layout.num_fields
        if len(self.fe_spaces) != nf:
            raise FemhipError(N.ERR_INVALID, "one FE space per field")
        L = N.BlockLayout()
        L.num_fields = nf
        for f in range(nf):
            L.fields[f] = self.fe_spaces[f].h.value if hasattr(self.fe_spaces[f].h, "value") else self.fe_spaces[f].h
            L.field_blocks[f] = block_layout.field_id_to_block_id[f]
            for g in range(nf):
                L.field_coupling[f * nf + g] = int(block_layout.field_coupling[f, g])
        env = self.fe_spaces[0].triangulation.environment
        h = C.c_void_p()
        N.check(self._lib.femhip_block_operator_create(env.h, C.byref(L), int(self.symmetric_storage), C.byref(h)))
        self.h = h
        nb = C.c_int32()
        N.check(self._lib.femhip_block_operator_info(h, C.byref(nb), None, None))
        self.num_blocks = nb.value
        sizes = np.zeros(self.num_blocks, dtype=np.int32)
        nnz = np.zeros(self.num_blocks * self.num_blocks, dtype=np.int64)
        N.check(self._lib.femhip_block_operator_info(h, C.byref(nb), N.ptr(sizes), N.ptr(nnz)))
        self.block_sizes = [int(v) for v in sizes]
        self.block_nnz = nnz.reshape(self.num_blocks, self.num_blocks)
        self.num_rows = int(sizes.sum())
        self._phys = discrete_integration.physics()
        N.check(self._lib.femhip_block_operator_set_physics(h, C.byref(self._phys)))
        self.last_compute_ms = None
        return self

    def compute(self):
        ms = C.c_double()
        N.check(self._lib.femhip_block_operator_compute(self.h, C.byref(ms)))
        self.last_compute_ms = ms.value

    def get_lst_dofs_gids(self, field_id):
        s = self.fe_spaces[field_id - 1]
        out = np.empty((s.num_cells, s.num_dofs_x_cell), dtype=np.int32)
        N.check(self._lib.femhip_block_operator_get_dofs(self.h, field_id, N.ptr(out)))
        return out

    def get_block(self, ib, jb) -> csr_matrix_view:
        """block_sparse_matrix_t%get_block(ib, jb) (1-based)."""
        n = self.block_sizes[ib - 1]
        nnz = int(self.block_nnz[ib - 1, jb - 1])
        irp = np.empty(n + 1, dtype=np.int32)
        ja = np.empty(max(nnz, 1), dtype=np.int32)
        val = np.empty(max(nnz, 1))
        N.check(self._lib.femhip_block_operator_get_block_pattern(self.h, ib, jb, N.ptr(irp), N.ptr(ja)))
        N.check(self._lib.femhip_block_operator_get_block_values(self.h, ib, jb, N.ptr(val)))
        return csr_matrix_view(irp, ja[:nnz], val[:nnz], self.symmetric_storage and ib == jb)

    def get_translation(self):
        """The block RHS, all blocks' rows in block order."""
        out = []
        for ib in range(1, self.num_blocks + 1):
            r = np.empty(max(self.block_sizes[ib - 1], 1))
            N.check(self._lib.femhip_block_operator_get_rhs(self.h, ib, N.ptr(r)))
            out.append(r[: self.block_sizes[ib - 1]])
        return np.concatenate(out)

    def get_tangent(self):
        return self

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.num_rows)
        N.check(self._lib.femhip_block_operator_apply(self.h, N.ptr(x), N.ptr(y)))
        return y

    def free(self):
        if getattr(self, "h", None):
            self._lib.femhip_block_operator_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class batch_t:
    """Many independent Poisson problems (Q1, unit square/cube, constant k,
    f = f0 k d pi^2 prod sin(pi x), homogeneous Dirichlet) assembled in one launch
    and solved in one persistent FP32-refined launch (femhip_batch_*)."""

    def __init__(self, environment: environment_t, problems, rtol=1e-8, max_num_iterations=20000, delta=0.0):
        self._lib = N.load()
        self.environment = environment
        arr = (N.BatchProblem * len(problems))()
        for i, (dim, n, k, f0) in enumerate(problems):
            arr[i].dim, arr[i].n, arr[i].k, arr[i].f0 = int(dim), int(n), float(k), float(f0)
        self.problems = list(problems)
        p = N.CgParams()
        p.rtol, p.atol, p.max_num_iterations = rtol, 0.0, max_num_iterations
        p.stopping_criterion = N.STOPPING_CRITERIA.index("RES_RES")
        p.preconditioner, p.precision, p.inner_rtol = N.PREC_JACOBI, N.PRECISION_FP32_REFINED, delta
        h = C.c_void_p()
        N.check(self._lib.femhip_batch_create(environment.h, len(problems), arr, C.byref(p), C.byref(h)))
        self.h = h

    def assemble(self):
        ms = C.c_double()
        N.check(self._lib.femhip_batch_assemble(self.h, C.byref(ms)))
        return ms.value

    def autotune(self):
        t = C.c_int32()
        N.check(self._lib.femhip_batch_autotune(self.h, C.byref(t)))
        return t.value

    def solve(self):
        res = (N.BatchResult * len(self.problems))()
        ms = C.c_double()
        N.check(self._lib.femhip_batch_solve(self.h, res, C.byref(ms)))
        out = {f: np.array([getattr(r, f) for r in res]) for f, _ in N.BatchResult._fields_}
        out["ms"] = ms.value
        return out

    def get_solution(self, i):
        dim, n = self.problems[i][0], self.problems[i][1]
        x = np.empty((n - 1) ** dim)
        N.check(self._lib.femhip_batch_get_solution(self.h, i, N.ptr(x)))
        return x

    def get_problem_matrix(self, i):
        dim, n = self.problems[i][0], self.problems[i][1]
        rows = (n - 1) ** dim
        cap = rows * (9 if dim == 2 else 27)
        irp = np.empty(rows + 1, dtype=np.int32)
        ja = np.empty(cap, dtype=np.int32)
        val = np.empty(cap)
        rhs = np.empty(rows)
        nnz = C.c_int64()
        N.check(self._lib.femhip_batch_get_problem_matrix(self.h, i, N.ptr(irp), N.ptr(ja), N.ptr(val), N.ptr(rhs),
                                                          C.byref(nnz)))
        return irp, ja[: nnz.value], val[: nnz.value], rhs

    def free(self):
        if getattr(self, "h", None):
            self._lib.femhip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class serve_t:
    """Request loop with continuous batching (femhip_serve_*): each request is
    a vector-Q1 linear elasticity solve on the unit cube with n cells per
    direction, clamped on x = 0, constant body force, FP32-refined Jacobi-PCG."""

    ADMISSION = {"fcfs": 0, "smallest": 1}

    def __init__(self, environment: environment_t, hbm_budget_bytes=64 << 30, max_active=8, setup_threads=4,
                 rtol=1e-8, max_num_iterations=20000, delta=0.0, admission="fcfs", aging_s=0.0):
        self._lib = N.load()
        self.environment = environment
        p = N.ServeParams()
        p.hbm_budget_bytes, p.max_active, p.setup_threads = float(hbm_budget_bytes), max_active, setup_threads
        p.admission, p.aging_s = self.ADMISSION[admission], float(aging_s)
        p.cg.rtol, p.cg.atol, p.cg.max_num_iterations = rtol, 0.0, max_num_iterations
        p.cg.stopping_criterion = N.STOPPING_CRITERIA.index("RES_RES")
        p.cg.preconditioner, p.cg.precision, p.cg.inner_rtol = N.PREC_JACOBI, N.PRECISION_FP32_REFINED, delta
        self.max_num_iterations = max_num_iterations
        h = C.c_void_p()
        N.check(self._lib.femhip_serve_create(environment.h, C.byref(p), C.byref(h)))
        self.h = h
        self._dofs = {}

    def estimate_bytes(self, n):
        return self._lib.femhip_serve_estimate_bytes(int(n), self.max_num_iterations)

    def now(self):
        return self._lib.femhip_serve_now(self.h)

    def submit(self, n, E, nu, body=(0.0, 0.0, -1.0), arrival_s=-1.0):
        rq = N.ServeRequest()
        rq.n, rq.E, rq.nu, rq.arrival_s = int(n), float(E), float(nu), float(arrival_s)
        for i in range(3):
            rq.body[i] = float(body[i])
        t = C.c_int64()
        N.check(self._lib.femhip_serve_submit(self.h, C.byref(rq), C.byref(t)))
        return t.value

    def wait(self, ticket, timeout_s=-1.0):
        r = N.ServeResult()
        N.check(self._lib.femhip_serve_wait(self.h, ticket, float(timeout_s), C.byref(r)))
        out = {f: getattr(r, f) for f, _ in N.ServeResult._fields_ if f != "pad"}
        out["error"] = self._lib.femhip_last_error().decode(errors="replace") if r.status not in (N.OK, 4) else ""
        self._dofs[ticket] = r.num_dofs
        return out

    def get_solution(self, ticket):
        x = np.empty(self._dofs[ticket])
        N.check(self._lib.femhip_serve_get_solution(self.h, ticket, N.ptr(x)))
        return x

    def release(self, ticket):
        N.check(self._lib.femhip_serve_release(self.h, ticket))
        self._dofs.pop(ticket, None)

    def stats(self):
        s = N.ServeStats()
        N.check(self._lib.femhip_serve_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in N.ServeStats._fields_}

    def free(self):
        if getattr(self, "h", None):
            self._lib.femhip_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
