"""fempar_amd -- MI355X-native finite-element hot path behind FEMPAR's API.

The compute path is libfemhip.so (HIP kernels for gfx950 behind the C-ABI of
include/femhip.h); this package holds its sources (csrc/), the build recipe and
a host-side mirror of FEMPAR's Fortran interfaces (fempar_names).
"""
from ._native import FemhipError, LIB_PATH  # noqa: F401
from .fempar_names import (  # noqa: F401
    environment_t, serial_triangulation_t, serial_fe_space_t, poisson_discrete_integration_t,
    linear_elasticity_discrete_integration_t, fe_affine_operator_t, iterative_linear_solver_t,
    jacobi_preconditioner_t, identity_operator_t, csr_matrix_view, batch_t, serve_t,
    block_layout_t, coupled_diffusion_discrete_integration_t, block_fe_affine_operator_t,
)
