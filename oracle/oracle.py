"""ctypes wrapper around the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, always as the checker.  Nothing in the
fempar_amd package may import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

BC_ALL, BC_X0, BC_NONE = 0, 1, 2
PHYS_POISSON, PHYS_ELASTICITY = 0, 1
K_CONST, K_SIN3, K_CELL, K_QPOINT = 0, 1, 2, 3
F_CONST, F_MANUFACTURED, F_CELL_SIN, F_QPOINT = 0, 1, 2, 3
G_ZERO, G_LINEAR = 0, 1
EXACT_SIN, EXACT_LINEAR = 0, 1

CRITERIA = ["RES_NRMGIVEN_RHS_NRMGIVEN", "RES_NRMGIVEN_RES_NRMGIVEN", "DELTA_RHS", "DELTA_DELTA",
            "RES_RES", "RES_RHS", "DELTA_RHS_AND_RES_RES", "DELTA_RHS_AND_RES_RHS",
            "DELTA_DELTA_AND_RES_RES", "DELTA_DELTA_AND_RES_RHS"]


class Problem(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int * 3), ("lim", C.c_double * 6), ("order", C.c_int),
                ("ncomp", C.c_int), ("bc_kind", C.c_int), ("vcoords", C.POINTER(C.c_double))]


class Physics(C.Structure):
    _fields_ = [("kind", C.c_int), ("k_kind", C.c_int), ("k0", C.c_double),
                ("k_cell", C.POINTER(C.c_double)), ("f_kind", C.c_int), ("f0", C.c_double),
                ("g_kind", C.c_int), ("E", C.c_double), ("nu", C.c_double), ("body", C.c_double * 3),
                ("k_qp", C.POINTER(C.c_double)), ("f_qp", C.POINTER(C.c_double))]


class CgResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("error_estimate", C.c_double),
                ("rhs_test", C.c_double)]


MF_MAXF = 4


class MFPhysics(C.Structure):
    _fields_ = [("kappa", C.c_double * MF_MAXF), ("sigma", C.c_double), ("f", C.c_double * MF_MAXF)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        L.orc_space_create.restype = P
        L.orc_space_create.argtypes = [C.POINTER(Problem)]
        L.orc_space_free.argtypes = [P]
        L.orc_space_sizes.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_space_get_dofs.argtypes = [P, P]
        L.orc_space_fixed_values.argtypes = [P, C.POINTER(Physics), P]
        L.orc_reference_tables.argtypes = [P, P, P, P, P]
        L.orc_gauss_legendre.argtypes = [C.c_int, P, P]
        L.orc_element.argtypes = [P, C.POINTER(Physics), C.c_int64, P, P]
        L.orc_quadrature_points.argtypes = [P, P]
        L.orc_assemble.restype = C.c_int64
        L.orc_assemble.argtypes = [P, C.POINTER(Physics), C.c_int, P, C.POINTER(P), C.POINTER(P),
                                   C.POINTER(P), C.POINTER(P)]
        L.orc_assemble_cells.restype = C.c_int64
        L.orc_assemble_cells.argtypes = [P, C.POINTER(Physics), C.c_int, P, C.c_int64, C.c_int64, C.POINTER(P),
                                         C.POINTER(P), C.POINTER(P), C.POINTER(P)]
        L.orc_free.argtypes = [P]
        L.orc_spmv.argtypes = [C.c_int, P, P, P, C.c_int, P, P]
        L.orc_extract_diagonal.argtypes = [C.c_int, P, P, P, P]
        L.orc_cg.argtypes = [C.c_int, P, P, P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                             C.c_int, P, P, C.POINTER(CgResult), P]
        L.orc_error_norms.argtypes = [P, P, P, C.c_int, P]
        L.orc_cg_threads.argtypes = [C.c_int, P, P, P, C.c_double, C.c_int, P, P, C.POINTER(CgResult)]
        L.orc_mf_create.restype = P
        L.orc_mf_create.argtypes = [C.c_int, P, P]
        L.orc_mf_free.argtypes = [P]
        L.orc_mf_num_blocks.argtypes = [P]
        L.orc_mf_field_sizes.argtypes = [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_mf_block_size.argtypes = [P, C.c_int]
        L.orc_mf_get_dofs.argtypes = [P, C.c_int, P]
        L.orc_mf_element.argtypes = [P, C.POINTER(MFPhysics), C.c_int64, P, P]
        L.orc_mf_assemble_block.restype = C.c_int64
        L.orc_mf_assemble_block.argtypes = [P, C.POINTER(MFPhysics), P, C.c_int, C.c_int, C.c_int, C.POINTER(P),
                                            C.POINTER(P), C.POINTER(P)]
        L.orc_mf_assemble_rhs.argtypes = [P, C.POINTER(MFPhysics), C.c_int, P]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_problem(dim, n, order=1, ncomp=1, bc=BC_ALL, lim=None, vcoords=None):
    """vcoords: optional (num_vertices, dim) vertex coordinates, vertex (i, j, k)
    at row i + (n0+1)(j + (n1+1) k); every cell then has its own map."""
    p = Problem()
    p._vc_ref = None
    if vcoords is not None:
        arr = np.ascontiguousarray(vcoords, dtype=np.float64).ravel()
        p._vc_ref = arr
        p.vcoords = arr.ctypes.data_as(C.POINTER(C.c_double))
    p.dim = dim
    nn = list(n) + [1] * (3 - len(n))
    for i in range(3):
        p.n[i] = nn[i]
    lim = lim or [0.0, 1.0, 0.0, 1.0, 0.0, 1.0 if dim == 3 else 0.0]
    for i in range(6):
        p.lim[i] = lim[i]
    p.order, p.ncomp, p.bc_kind = order, ncomp, bc
    return p


def make_physics(kind=PHYS_POISSON, k_kind=K_CONST, k0=1.0, k_cell=None, f_kind=F_CONST, f0=0.0,
                 g_kind=G_ZERO, E=1.0, nu=0.3, body=(0.0, 0.0, -1.0), k_qp=None, f_qp=None):
    ph = Physics()
    ph.kind, ph.k_kind, ph.k0 = kind, k_kind, k0
    ph._refs = []
    for name, a in (("k_cell", k_cell), ("k_qp", k_qp), ("f_qp", f_qp)):
        if a is not None:
            arr = np.ascontiguousarray(a, dtype=np.float64).ravel()
            ph._refs.append(arr)
            setattr(ph, name, arr.ctypes.data_as(C.POINTER(C.c_double)))
    ph.f_kind, ph.f0, ph.g_kind, ph.E, ph.nu = f_kind, f0, g_kind, E, nu
    for i in range(3):
        ph.body[i] = body[i]
    return ph


class Space:
    """Structured mesh + reference FE + DOF numbering (oracle)."""

    def __init__(self, problem: Problem):
        self.problem = problem
        self.h = lib().orc_space_create(C.byref(problem))
        if not self.h:
            raise ValueError("oracle: unsupported problem")
        nc, nd, nq, nf, nx = C.c_int64(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib().orc_space_sizes(self.h, C.byref(nc), C.byref(nd), C.byref(nq), C.byref(nf), C.byref(nx))
        self.num_cells, self.nd, self.nq = nc.value, nd.value, nq.value
        self.n_free, self.n_fixed = nf.value, nx.value

    def __del__(self):
        try:
            if self.h:
                lib().orc_space_free(self.h)
        except Exception:
            pass

    def dofs(self):
        out = np.empty((self.num_cells, self.nd), dtype=np.int32)
        lib().orc_space_get_dofs(self.h, _ptr(out))
        return out

    def fixed_values(self, phys):
        out = np.zeros(max(self.n_fixed, 1), dtype=np.float64)
        lib().orc_space_fixed_values(self.h, C.byref(phys), _ptr(out))
        return out[: self.n_fixed]

    def tables(self):
        d = self.problem.dim
        nnode = self.nd // self.problem.ncomp
        qp = np.empty((self.nq, d)); qw = np.empty(self.nq)
        shp = np.empty((self.nq, nnode)); dshp = np.empty((self.nq, nnode, d))
        lib().orc_reference_tables(self.h, _ptr(qp), _ptr(qw), _ptr(shp), _ptr(dshp))
        return qp, qw, shp, dshp

    def error_norms(self, x_free, fixed_values=None, exact="sin"):
        """l2_norm, h1_seminorm, h1_norm, linfty_norm of u - u_h
        (error_norms_scalar_t%compute) for u = prod sin(pi x_d) or x + y (+ z)."""
        x = np.ascontiguousarray(x_free, dtype=np.float64)
        fx = None if fixed_values is None else np.ascontiguousarray(fixed_values, dtype=np.float64)
        out = np.zeros(4)
        lib().orc_error_norms(self.h, _ptr(x), _ptr(fx), {"sin": EXACT_SIN, "linear": EXACT_LINEAR}[exact],
                              _ptr(out))
        return {"l2_norm": out[0], "h1_seminorm": out[1], "h1_norm": out[2], "linfty_norm": out[3]}

    def quadrature_points(self):
        """Physical coordinates of every quadrature point, (num_cells, nq, dim)."""
        d = self.problem.dim
        out = np.empty((self.num_cells, self.nq, d), dtype=np.float64)
        lib().orc_quadrature_points(self.h, _ptr(out))
        return out

    def element(self, phys, cell):
        em = np.empty((self.nd, self.nd), dtype=np.float64, order="F")
        ev = np.empty(self.nd)
        lib().orc_element(self.h, C.byref(phys), cell, em.ctypes.data_as(C.c_void_p), _ptr(ev))
        return em, ev

    def assemble(self, phys, symmetric_storage=True, fixed_values=None, cells=None):
        irp, ja, val, rhs = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        fv = None if fixed_values is None else np.ascontiguousarray(fixed_values, dtype=np.float64)
        cb, ce = (0, self.num_cells) if cells is None else cells
        nnz = lib().orc_assemble_cells(self.h, C.byref(phys), int(bool(symmetric_storage)), _ptr(fv), cb, ce,
                                       C.byref(irp), C.byref(ja), C.byref(val), C.byref(rhs))
        n = self.n_free

        def take(ptr, count, ctype, dtype):
            if count == 0:
                lib().orc_free(ptr)
                return np.zeros(0, dtype=dtype)
            arr = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(count,)).copy()
            lib().orc_free(ptr)
            return arr.astype(dtype, copy=False)

        return (take(irp, n + 1, C.c_int32, np.int32), take(ja, nnz, C.c_int32, np.int32),
                take(val, nnz, C.c_double, np.float64), take(rhs, n, C.c_double, np.float64))


def gauss_legendre(n):
    x = np.empty(n); w = np.empty(n)
    lib().orc_gauss_legendre(n, _ptr(x), _ptr(w))
    return x, w


def spmv(irp, ja, val, x, symmetric_storage):
    n = len(irp) - 1
    y = np.empty(n)
    lib().orc_spmv(n, _ptr(irp), _ptr(ja), _ptr(val), int(bool(symmetric_storage)),
                   _ptr(np.ascontiguousarray(x, dtype=np.float64)), _ptr(y))
    return y


def extract_diagonal(irp, ja, val):
    n = len(irp) - 1
    d = np.empty(n)
    lib().orc_extract_diagonal(n, _ptr(irp), _ptr(ja), _ptr(val), _ptr(d))
    return d


def cg(irp, ja, val, b, symmetric_storage=True, precond="jacobi", criterion="RES_RES", rtol=1e-10,
       atol=0.0, maxit=5000, x0=None, history=False):
    n = len(irp) - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    res = CgResult()
    hist = np.zeros(maxit) if history else None
    lib().orc_cg(n, _ptr(irp), _ptr(ja), _ptr(val), int(bool(symmetric_storage)),
                 1 if precond == "jacobi" else 0, CRITERIA.index(criterion), rtol, atol, maxit,
                 _ptr(np.ascontiguousarray(b, dtype=np.float64)), _ptr(x), C.byref(res), _ptr(hist))
    out = {"iterations": res.iterations, "converged": bool(res.converged),
           "error_estimate": res.error_estimate, "rhs_test": res.rhs_test}
    if history:
        out["history"] = hist[: res.iterations].copy()
    return x, out


def _take(ptr, count, ctype, dtype):
    if count == 0:
        lib().orc_free(ptr)
        return np.zeros(0, dtype=dtype)
    arr = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(count,)).copy()
    lib().orc_free(ptr)
    return arr.astype(dtype, copy=False)


def make_mf_physics(kappa, sigma, f):
    ph = MFPhysics()
    for i in range(MF_MAXF):
        ph.kappa[i] = kappa[i] if i < len(kappa) else 0.0
        ph.f[i] = f[i] if i < len(f) else 0.0
    ph.sigma = sigma
    return ph


class MultiFieldSpace:
    """Several scalar fields on one structured mesh with a block layout
    (oracle): per-field numbering with block offsets, block assembly."""

    def __init__(self, problems, field_blocks):
        self.nf = len(problems)
        arr = (Problem * self.nf)(*problems)
        fb = np.ascontiguousarray(field_blocks, dtype=np.int32)
        self._keep = (arr, fb)
        self.h = lib().orc_mf_create(self.nf, C.cast(arr, C.c_void_p), _ptr(fb))
        if not self.h:
            raise ValueError("oracle: unsupported multi-field problem")
        self.field_blocks = list(field_blocks)
        self.num_blocks = lib().orc_mf_num_blocks(self.h)
        self.nd, self.n_free, self.n_fixed = [], [], []
        for f in range(self.nf):
            a, b, c = C.c_int(), C.c_int(), C.c_int()
            lib().orc_mf_field_sizes(self.h, f, C.byref(a), C.byref(b), C.byref(c))
            self.nd.append(a.value); self.n_free.append(b.value); self.n_fixed.append(c.value)
        self.block_sizes = [lib().orc_mf_block_size(self.h, b) for b in range(1, self.num_blocks + 1)]
        self.num_cells = 1
        for k in range(problems[0].dim):
            self.num_cells *= problems[0].n[k]

    def __del__(self):
        try:
            if self.h:
                lib().orc_mf_free(self.h)
        except Exception:
            pass

    def dofs(self, field):
        out = np.empty((self.num_cells, self.nd[field]), dtype=np.int32)
        lib().orc_mf_get_dofs(self.h, field, _ptr(out))
        return out

    def element(self, phys, cell=0):
        ndt = sum(self.nd)
        em = np.empty((ndt, ndt), dtype=np.float64, order="F")
        ev = np.empty(ndt)
        lib().orc_mf_element(self.h, C.byref(phys), cell, em.ctypes.data_as(C.c_void_p), _ptr(ev))
        return em, ev

    def assemble_block(self, phys, coupling, ib, jb, symmetric_storage=False):
        """Block (ib, jb), 1-based: irp, ja (1-based), val."""
        cp = np.ascontiguousarray(np.asarray(coupling, dtype=bool).astype(np.int32).ravel())
        irp, ja, val = C.c_void_p(), C.c_void_p(), C.c_void_p()
        nnz = lib().orc_mf_assemble_block(self.h, C.byref(phys), _ptr(cp), ib, jb, int(bool(symmetric_storage)),
                                          C.byref(irp), C.byref(ja), C.byref(val))
        n = self.block_sizes[ib - 1]
        return (_take(irp, n + 1, C.c_int32, np.int32), _take(ja, nnz, C.c_int32, np.int32),
                _take(val, nnz, C.c_double, np.float64))

    def assemble_rhs(self, phys, ib):
        out = np.zeros(max(1, self.block_sizes[ib - 1]))
        lib().orc_mf_assemble_rhs(self.h, C.byref(phys), ib, _ptr(out))
        return out[: self.block_sizes[ib - 1]]

    def merged(self, phys, coupling):
        """The block system as one CSR (block rows in order, block columns
        shifted by the sizes of the earlier blocks) and its RHS."""
        offs = np.concatenate([[0], np.cumsum(self.block_sizes)]).astype(np.int64)
        rows = []
        for ib in range(1, self.num_blocks + 1):
            parts = [self.assemble_block(phys, coupling, ib, jb) for jb in range(1, self.num_blocks + 1)]
            for r in range(self.block_sizes[ib - 1]):
                cols, vals = [], []
                for jb, (irp, ja, val) in enumerate(parts):
                    a, b = irp[r] - 1, irp[r + 1] - 1
                    cols.append(ja[a:b] + offs[jb]); vals.append(val[a:b])
                rows.append((np.concatenate(cols), np.concatenate(vals)))
        irp = np.zeros(len(rows) + 1, dtype=np.int32)
        irp[1:] = np.cumsum([len(c) for c, _ in rows])
        irp += 1
        ja = np.concatenate([c for c, _ in rows]).astype(np.int32)
        val = np.concatenate([v for _, v in rows])
        rhs = np.concatenate([self.assemble_rhs(phys, ib) for ib in range(1, self.num_blocks + 1)])
        return irp, ja, val, rhs


def cg_threads(irp, ja, val, b, rtol=1e-10, maxit=5000):
    """Jacobi-PCG (RES_RES) on a full-storage CSR over the host threads: the
    all-cores timing baseline (not the serial reference's algorithm)."""
    n = len(irp) - 1
    x = np.zeros(n)
    res = CgResult()
    lib().orc_cg_threads(n, _ptr(irp), _ptr(ja), _ptr(val), rtol, maxit,
                         _ptr(np.ascontiguousarray(b, dtype=np.float64)), _ptr(x), C.byref(res))
    return x, {"iterations": res.iterations, "converged": bool(res.converged)}
