"""The reference's 1.5-layer QG double-gyre example on the GPU
(``scripts/qg_1p5_layer.py``), composed from the drop-in operators.

Each step is ``heun_step`` of

    dq/dt = -div(u q) - beta v + F - r zeta + nu lap(q),
    (lap - 1/Ld^2) psi = q  (DST-I, psi = 0 on the walls),

with upwind1 advection, u = -d_y(T_to_X psi), v = d_x(T_to_X psi) at the
U / V faces, zeta = lap(psi) and the wall (zero ghost ring) condition
re-applied before each tendency and after the step (``:420-497``).  Every
operator is one HIP kernel of ``libfvx_hip.so``; the Helmholtz inversion is the
hipFFT DST-I solve.  Snapshots are returned as arrays (the reference writes
them with xarray / Zarr, absent here).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from finitevolx_amd import _lib
from finitevolx_amd.advection import Advection2D
from finitevolx_amd.boundary import pad_interior
from finitevolx_amd.grid import CartesianGrid2D
from finitevolx_amd.operators import Difference2D, Interpolation2D
from finitevolx_amd.solvers import solve_helmholtz_dst
from finitevolx_amd.timestepping import heun_step, lincomb


@dataclass
class QuasiGeostrophicConfig:
    """Defaults of ``QuasiGeostrophicConfig`` (``scripts/qg_1p5_layer.py:280-300``)."""

    nx: int = 64
    ny: int = 64
    Lx: float = 5.12e6
    Ly: float = 5.12e6
    f0: float = 9.375e-5
    beta: float = 1.754e-11
    gravity: float = 9.81
    rossby_radius: float = 4.0e4
    drag: float = 5.0e-8
    viscosity: float = 5.0e4
    wind_curl_forcing: float = 2.0e-12
    dt: float = 4000.0
    spinup_steps: int = 5000
    steps: int = 10000
    snapshot_interval: int = 1000


def to_wall_field(interior) -> torch.Tensor:
    """Interior ``[ny, nx]`` -> full ``[ny+2, nx+2]`` with a zero ghost ring
    (``scripts/qg_1p5_layer.py:92-113``)."""
    t = _lib.as_field(interior)
    return torch.nn.functional.pad(t[None, None], (1, 1, 1, 1))[0, 0].contiguous()


def geostrophic_velocity_from_streamfunction(psi_field, diff: Difference2D,
                                             interp: Interpolation2D):
    """T-point psi -> (u, v) at U / V faces through X points (``:115-145``)."""
    psi_on_x = interp.T_to_X(psi_field)
    return -diff.diff_y_X_to_U(psi_on_x), diff.diff_x_X_to_V(psi_on_x)


def initial_fields(config: QuasiGeostrophicConfig, grid: CartesianGrid2D):
    """PV anomaly, beta term and wind-curl forcing on the interior
    (``make_preprocessing_dataset``, ``:302-371``), as host arrays."""
    x = (np.arange(config.nx) + 0.5) * grid.dx
    y = (np.arange(config.ny) + 0.5) * grid.dy
    X, Y = np.meshgrid(x, y)
    q0 = 5.0e-9 * np.sin(2.0 * np.pi * X / config.Lx) * np.sin(np.pi * Y / config.Ly)
    beta_term = config.beta * (Y - 0.5 * config.Ly)
    wind_curl = -config.wind_curl_forcing * np.sin(2.0 * np.pi * Y / config.Ly)
    return q0, beta_term, wind_curl


class QG1p5LayerModel:
    """The example's state update on device arrays (FP64 by default)."""

    def __init__(self, config: QuasiGeostrophicConfig | None = None,
                 dtype: torch.dtype = torch.float64):
        self.config = config or QuasiGeostrophicConfig()
        c = self.config
        self.grid = CartesianGrid2D.from_interior(c.nx, c.ny, c.Lx, c.Ly)
        self.diff = Difference2D(self.grid)
        self.interp = Interpolation2D(self.grid)
        self.adv = Advection2D(self.grid)
        self.dtype = dtype
        q0, _, wind = initial_fields(c, self.grid)
        self.q0 = to_wall_field(torch.tensor(q0, dtype=dtype))
        self.wind_curl = to_wall_field(torch.tensor(wind, dtype=dtype))
        self.lam = 1.0 / c.rossby_radius**2

    def invert_streamfunction(self, q_field):
        """(lap - 1/Ld^2) psi = q on the interior, psi = 0 walls (``:432-452``)."""
        psi_i = solve_helmholtz_dst(q_field[1:-1, 1:-1].contiguous(), self.grid.dx,
                                    self.grid.dy, self.lam)
        return to_wall_field(psi_i)

    def tendency(self, q_field):
        """``(q_rhs, psi, u, v)`` of ``tendency`` (``:454-478``)."""
        c = self.config
        psi = self.invert_streamfunction(q_field)
        u, v = geostrophic_velocity_from_streamfunction(psi, self.diff, self.interp)
        q_rhs = self.adv(q_field, u, v, method="upwind1")
        # V_to_T's ghost ring is zero, so the full-array combination adds
        # -beta v on the interior only, as the reference's ``.at[1:-1, 1:-1]``
        v_center = self.interp.V_to_T(v)
        q_rhs = lincomb(1.0, q_rhs, -c.beta, v_center, 1.0, self.wind_curl)
        zeta = self.diff.laplacian(psi)
        q_rhs = lincomb(1.0, q_rhs, -c.drag, zeta, c.viscosity, self.diff.laplacian(q_field))
        return q_rhs, psi, u, v

    @staticmethod
    def apply_bc(q_field):
        """Zero ghost ring (``pad_interior``, ``:480-482``)."""
        return pad_interior(q_field, mode="constant")

    def pv_tendency(self, q_field):
        return self.tendency(self.apply_bc(q_field))[0]

    def step(self, q_field):
        """One Heun step and the diagnosed balanced state (``:484-497``)."""
        q_next = self.apply_bc(heun_step(q_field, self.pv_tendency, self.config.dt))
        psi_next = self.invert_streamfunction(q_next)
        u_next, v_next = geostrophic_velocity_from_streamfunction(psi_next, self.diff,
                                                                  self.interp)
        return q_next, psi_next, u_next, v_next

    def run(self, steps: int | None = None, snapshot_interval: int | None = None):
        """Integrate from the initial PV; returns the snapshot times and the
        q, psi, u, v snapshots as host arrays."""
        c = self.config
        steps = c.steps if steps is None else steps
        every = c.snapshot_interval if snapshot_interval is None else snapshot_interval
        q = self.q0.clone()
        out = {"time": [], "q": [], "psi": [], "u": [], "v": []}
        for n in range(1, steps + 1):
            q, psi, u, v = self.step(q)
            if n % every == 0 or n == steps:
                out["time"].append(n * c.dt)
                for k, t in (("q", q), ("psi", psi), ("u", u), ("v", v)):
                    out[k].append(t.cpu().numpy())
        return {k: np.asarray(v) for k, v in out.items()}
