"""``import finitevolx`` for the MI355X build: the reference package's public
names for the quasi-geostrophic hot path (advection, Arakawa Jacobian, C-grid
stencils, diffusion, multigrid / CG / spectral Helmholtz solves, vertical
modes, explicit Runge-Kutta stepping), each backed by the HIP library through
``finitevolx_amd``.  Names outside that path are not provided."""

from finitevolx_amd import *  # noqa: F401,F403
from finitevolx_amd import __all__  # noqa: F401
