"""finitevolx_amd: MI355X-native (gfx950 HIP) QG hot path of finitevolX.

Same public names as the reference's ``finitevolx`` package for the
quasi-geostrophic step: advection, Arakawa Jacobian, C-grid stencils,
diffusion, multigrid-preconditioned Helmholtz solves, vertical modes and
SSP-RK3 stepping.  All compute runs in ``libfvx_hip.so``; there is no CPU
fallback.
"""

from finitevolx_amd.advection import Advection2D, Reconstruction2D
from finitevolx_amd.boundary import enforce_periodic, pad_interior
from finitevolx_amd.diffusion import (
    BiharmonicDiffusion2D,
    Diffusion2D,
    MomentumAdvection2D,
    diffusion_2d,
)
from finitevolx_amd.grid import ArakawaCGrid2D, CartesianGrid2D
from finitevolx_amd.mask import Mask2D, StencilCapability2D
from finitevolx_amd.operators import (
    Difference2D,
    Divergence2D,
    Interpolation2D,
    arakawa_jacobian,
    divergence_2d,
    interior,
)
from finitevolx_amd.qg import QGModel
from finitevolx_amd.solvers import (
    CGInfo,
    MultigridLevel,
    MultigridSolver,
    build_multigrid_solver,
    make_multigrid_preconditioner,
    masked_laplacian,
    pv_inversion,
    dst1_eigenvalues,
    solve_cg,
    solve_helmholtz_dct,
    solve_helmholtz_dst,
    solve_helmholtz_fft,
    solve_poisson_dct,
    solve_poisson_dst,
    solve_poisson_fft,
    streamfunction_from_vorticity,
)
from finitevolx_amd.timestepping import euler_step, heun_step, rk3_ssp_step, rk4_step
from finitevolx_amd.vertical import (
    build_coupling_matrix,
    decompose_vertical_modes,
    layer_to_mode,
    mode_to_layer,
)

__all__ = [
    "Advection2D", "Reconstruction2D", "ArakawaCGrid2D", "BiharmonicDiffusion2D", "CGInfo", "CartesianGrid2D",
    "Difference2D", "Diffusion2D", "MomentumAdvection2D", "Divergence2D", "Interpolation2D", "Mask2D",
    "MultigridLevel", "MultigridSolver", "QGModel", "StencilCapability2D", "arakawa_jacobian",
    "build_coupling_matrix", "build_multigrid_solver", "decompose_vertical_modes",
    "diffusion_2d", "divergence_2d", "enforce_periodic", "interior", "layer_to_mode",
    "make_multigrid_preconditioner", "masked_laplacian", "mode_to_layer", "pad_interior",
    "pv_inversion", "euler_step", "heun_step", "rk3_ssp_step", "rk4_step", "solve_cg", "streamfunction_from_vorticity",
    "dst1_eigenvalues", "solve_helmholtz_dct", "solve_helmholtz_dst", "solve_helmholtz_fft",
    "solve_poisson_dct", "solve_poisson_dst", "solve_poisson_fft",
]
