"""FEMPAR's test_poisson sequence written in Fortran, run on the GPU: u = x + y
(+ z) must be reproduced at the nodes (test_poisson_driver.f90:562-600).

test_poisson_hip calls the ISO_C_BINDING module directly; test_poisson_seams
calls only the reference's plug-in types (fempar_amd/fortran/fempar_seams.f90:
discrete_integration_t, assembler_t, fe_affine_operator_t,
base_iterative_linear_solver_t, the creational-methods dictionary) with the HIP
subtypes of hip_fempar_plugins.f90 installed at start-up."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fempar_amd")
EXE = os.path.join(PKG, "test_poisson_hip")
SEAMS = os.path.join(PKG, "test_poisson_seams")


@pytest.mark.parametrize("dim,n,order", [(3, 8, 1), (2, 10, 2), (3, 4, 2)])
def test_fortran_test_poisson(dim, n, order):
    assert os.path.exists(EXE), "build the Fortran driver first (fempar_amd.build.build_fortran)"
    p = subprocess.run([EXE, str(dim), str(n), str(order)], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0 and "PASSED" in p.stdout


@pytest.mark.parametrize("dim,n,order,prec", [(3, 8, 1, 0), (2, 10, 2, 1), (3, 4, 2, 0), (2, 16, 1, 1)])
def test_fortran_test_poisson_through_the_plugin_seams(dim, n, order, prec):
    assert os.path.exists(SEAMS), "build the Fortran drivers first (fempar_amd.build.build_fortran)"
    p = subprocess.run([SEAMS, str(dim), str(n), str(order), str(prec)], capture_output=True, text=True,
                       timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0 and "PASSED" in p.stdout


@pytest.mark.parametrize("dim,n,order,prec", [(3, 6, 1, 1), (2, 12, 2, 1), (3, 4, 2, 0)])
def test_fortran_seams_moved_vertices_and_source_term(dim, n, order, prec):
    """The same driver on a mesh with moved interior vertices, the plug-in
    evaluating its scalar_function_t source term at every quadrature point
    (femhip_space_quadrature_points, then FEMHIP_F_QPOINT values); the nodes
    of the check are the mapped ones."""
    assert os.path.exists(SEAMS), "build the Fortran drivers first (fempar_amd.build.build_fortran)"
    p = subprocess.run([SEAMS, str(dim), str(n), str(order), str(prec), "1"], capture_output=True, text=True,
                       timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0 and "PASSED" in p.stdout and "moved=1" in p.stdout
