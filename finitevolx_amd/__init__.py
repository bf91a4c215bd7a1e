"""finitevolx_amd: MI355X-native (gfx950 HIP) QG hot path of finitevolX.

Same public names as the reference's ``finitevolx`` package for the
quasi-geostrophic step: advection, Arakawa Jacobian, C-grid stencils,
diffusion, multigrid-preconditioned Helmholtz solves and SSP-RK3 stepping.
All compute runs in ``libfvx_hip.so``; there is no CPU fallback.
"""

from finitevolx_amd.advection import Advection2D
from finitevolx_amd.boundary import enforce_periodic, pad_interior
from finitevolx_amd.diffusion import BiharmonicDiffusion2D, Diffusion2D, diffusion_2d
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
from finitevolx_amd.timestepping import rk3_ssp_step

__all__ = [
    "Advection2D", "ArakawaCGrid2D", "BiharmonicDiffusion2D", "CartesianGrid2D",
    "Difference2D", "Diffusion2D", "Divergence2D", "Interpolation2D", "Mask2D",
    "StencilCapability2D", "arakawa_jacobian", "diffusion_2d", "divergence_2d",
    "enforce_periodic", "interior", "pad_interior", "rk3_ssp_step",
]
