"""FEMPAR's test_poisson sequence written in Fortran against the ISO_C_BINDING
module (fempar_amd/fortran), run on the GPU: u = x + y (+ z) must be
reproduced at the nodes (test_poisson_driver.f90:562-600)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fempar_amd", "test_poisson_hip")


@pytest.mark.parametrize("dim,n,order", [(3, 8, 1), (2, 10, 2), (3, 4, 2)])
def test_fortran_test_poisson(dim, n, order):
    assert os.path.exists(EXE), "build the Fortran driver first (fempar_amd.build.build_fortran)"
    p = subprocess.run([EXE, str(dim), str(n), str(order)], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0 and "PASSED" in p.stdout
