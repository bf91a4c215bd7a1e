"""Multi-layer QG in a closed, masked basin with a passive tracer.

The layers couple only through the streamfunction inversion, which the
reference's vertical-mode API diagonalises (``vertical/vertical_modes.py``):
``radii, Cl2m, Cm2l = decompose_vertical_modes(build_coupling_matrix(H, g'),
f0)``; mode k solves ``(lap - 1/radii_k^2) psi_k = (Cl2m q)_k`` with psi = 0
on land and the walls (an infinite radius, from a non-positive eigenvalue of
the symmetrised coupling matrix, gives a Poisson solve), and the layer
streamfunctions are ``Cm2l psi_modes``.  Layer l then steps as the
1.5-layer model, ``dq_l/dt = -J(psi_l, q_l + beta (y - y0)) + F delta_{l0}
- nu4 D(D(q_l))``, with the wind forcing on the top layer only.  A tracer c
rides on the top layer's geostrophic face velocities
(``scripts/qg_1p5_layer.py:115-145``) with masked flux-form Advection2D
(WENO5 by default).  PV and tracer advance together by ``rk3_ssp_step``.

Each RK stage is a short fixed sequence of launches, replayed from a
captured HIP graph: the layer -> mode transform of the stage PV (one pass
over all layers), the extrapolated warm start of the modal streamfunctions,
the modal solves as one ``MultigridBatch`` (every operator of the solve one
launch over all modes; FP32 V-cycles inside FP64 iterative refinement,
certified at ``rtol`` in FP64), the mode -> layer transform, the tracer face
velocities, the QG stage of every layer (one work-list launch, FP64 state,
FP32 tendency) and the tracer's fused tendency + stage combination.
Solve budgets per stage are calibrated, tightened and widened as in
``finitevolx_amd.ensemble``.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from finitevolx_amd import _lib
from finitevolx_amd.advection import _MASK_DISPATCHABLE, method_codes
from finitevolx_amd.ensemble import (COEF, BudgetedSteps, _StageDesc, extrapolated_guess,
                                     hierarchy_depth, stage_psi_buffer)
from finitevolx_amd.mask import Mask2D
from finitevolx_amd.solvers import MultigridBatch, build_multigrid_solver
from finitevolx_amd.vertical import build_coupling_matrix, decompose_vertical_modes

_vp, _i, _d = _lib._vp, _lib._i, _lib._d

_lib.extra_signatures({
    **{f"fvx_advection_stage_{s}": [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d,
                                    _d, _d, _d, _vp] for s in ("f32", "f64")},
    **{f"fvx_geostrophic_faces_{s}": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _vp]
       for s in ("f32", "f64")},
})


class MultiLayerQG(BudgetedSteps):
    """Mixed-precision multi-layer QG steps with a passive tracer.

    ``q0``: initial layer PV, (nl, ny, nx) interior (layer 0 on top);
    ``c0``: initial tracer, (ny, nx) interior; ``mask``: optional (ny, nx)
    wet mask of the basin interior (the outer ring is always wall);
    ``chebyshev``: the lower end lo of [lo, 2] whose Chebyshev weights the
    modal solves' V-cycle sweeps take, one value or one per vertical mode
    (None: the uniform omega).  Measured on 3 layers at 4096^2: 29.6 steps/s
    uniform, 39.8 / 39.0 / 38.1 at lo = 0.15 / 0.2 / 0.25 (5.60 -> 4.18 inner
    iterations per modal solve); per mode, (0.15, 0.3, 0.3) 42.1 against
    41.0 at 0.2 on one box (within 3%, so one value stays the default).
    """

    def __init__(self, nx: int, ny: int, Lx: float, Ly: float, H, g_prime, q0, c0=None, *,
                 mask=None, f0: float = 9.375e-5, beta: float = 1.754e-11, nu4: float = 1e9,
                 F0: float = 2e-12, dt: float = 900.0, tracer_dtype=torch.float32,
                 tracer_method: str = "weno5", n_pre: int = 3, n_post: int = 3,
                 n_coarse: int = 4, omega: float = 0.85, rtol: float = 1e-8,
                 inner_rtol: float = 1e-4, chebyshev: float | None = 0.2):
        dev = _lib.require_gpu()
        self.nx, self.ny, self.Lx, self.Ly = nx, ny, Lx, Ly
        self.dx, self.dy = Lx / nx, Ly / ny
        self.f0, self.beta, self.nu4, self.F0, self.dt = f0, beta, nu4, F0, dt
        self.rtol, self.inner_rtol = rtol, inner_rtol
        A = build_coupling_matrix(H, g_prime)
        self.radii, self.Cl2m, self.Cm2l = decompose_vertical_modes(A, f0)
        self.mode_lambda = np.where(np.isfinite(self.radii), 1.0 / self.radii**2, 0.0)
        nl = self.nl = len(self.radii)
        if nl > 8:
            raise ValueError("at most 8 layers")
        Ny, Nx = ny + 2, nx + 2
        self.n_cells = Ny * Nx
        wet = np.ones((ny, nx)) if mask is None else (np.asarray(mask) > 0).astype(np.float64)
        if wet.shape != (ny, nx):
            raise ValueError(f"mask shape {wet.shape} != interior {(ny, nx)}")
        self.wet = wet
        full = np.zeros((Ny, Nx), dtype=bool)
        full[1:-1, 1:-1] = wet > 0
        self.mask2d = Mask2D(full)
        self.mask_h = self.mask2d.device("h")

        z = lambda: torch.zeros(nl, Ny, Nx, dtype=torch.float64, device=dev)  # noqa: E731
        self.q, self.q1, self.q2, self.backup = z(), z(), z(), z()
        self.qm = z()       # modal PV of the current stage
        self.psi_l = z()    # layer streamfunctions of the current stage
        self.ps = [z() for _ in range(4)]  # modal streamfunction history
        q0 = np.asarray(q0, dtype=np.float64)
        if q0.shape != (nl, ny, nx):
            raise ValueError(f"q0 must be {(nl, ny, nx)}, got {q0.shape}")
        self.q[:, 1:-1, 1:-1] = torch.from_numpy(q0 * wet).to(dev)

        self.tracer_method = tracer_method
        self.tm, self.tlim = method_codes(tracer_method)
        td = tracer_dtype
        zc = lambda: torch.zeros(Ny, Nx, dtype=td, device=dev)  # noqa: E731
        self.c, self.c1, self.c2, self.cbackup, self.u, self.v = (zc() for _ in range(6))
        if c0 is not None:
            c0 = np.asarray(c0, dtype=np.float64)
            if c0.shape != (ny, nx):
                raise ValueError(f"c0 must be {(ny, nx)}, got {c0.shape}")
            self.c[1:-1, 1:-1] = torch.from_numpy(c0 * wet).to(device=dev, dtype=td)
        self._csuf = _lib.suffix(self.c)
        masked = mask is not None
        self._vel_masks = ((self.mask2d.device("xy_corner_strict"), self.mask2d.device("u"),
                            self.mask2d.device("v")) if masked else (None, None, None))
        self._tiers = ((self.mask2d.device("tier_x"), self.mask2d.device("tier_y"))
                       if masked and tracer_method in _MASK_DISPATCHABLE else (None, None))
        self._mh = self.mask_h if masked else None

        cheb = list(chebyshev) if np.ndim(chebyshev) else [chebyshev] * nl
        if len(cheb) != nl:
            raise ValueError(f"chebyshev: one value or one per vertical mode ({nl})")
        self.solvers = [build_multigrid_solver(
            wet, self.dx, self.dy, float(self.mode_lambda[k]),
            n_levels=hierarchy_depth(ny, nx, min(self.dx, self.dy), float(self.radii[k])),
            n_pre=n_pre, n_post=n_post, n_coarse=n_coarse, omega=omega, kind="dirichlet",
            precision="mixed", chebyshev=cheb[k]) for k in range(nl)]
        self.batch = MultigridBatch(self.solvers)
        self._stage = {}
        self._init_budgets()

    def __del__(self):
        for h in getattr(self, "_stage", {}).values():
            try:
                _lib.call("fvx_qg_stage_batch_destroy", h)
            except Exception:
                pass
        self._stage = {}

    # -- state -----------------------------------------------------------
    @property
    def psi(self):
        """Layer streamfunctions of the last stage, (nl, ny+2, nx+2)."""
        return self.psi_l

    def modal_psi(self):
        """Modal streamfunctions of the last stage."""
        return self.ps[2]

    # -- launches --------------------------------------------------------
    def _stage_list(self, k):
        if k not in self._stage:
            qs = (self.q, self.q1, self.q2)[k]
            out = (self.q1, self.q2, self.q)[k]
            a, b, c = COEF[k]
            descs = (_StageDesc * self.nl)()
            for l in range(self.nl):  # noqa: E741
                d = descs[l]
                d.qs, d.psi = qs[l].data_ptr(), self.psi_l[l].data_ptr()
                d.qn, d.out = self.q[l].data_ptr(), out[l].data_ptr()
                d.mask = self.mask_h.data_ptr()
                d.Ny, d.Nx = self.ny + 2, self.nx + 2
                d.dx, d.dy = self.dx, self.dy
                d.beta, d.Ly, d.nu4 = self.beta, self.Ly, self.nu4
                d.F0 = self.F0 if l == 0 else 0.0
                d.a, d.b, d.cdt = a, b, c * self.dt
            h = _vp()
            _lib.call("fvx_qg_stage_batch_create", self.nl, descs, 1, C.byref(h))
            self._stage[k] = h
        return self._stage[k]

    def _mix(self, src, dst, M):
        Mc = np.ascontiguousarray(M, dtype=np.float64)
        _lib.call("fvx_layer_mix_f64", src.data_ptr(), dst.data_ptr(), Mc.ctypes.data, self.nl,
                  self.n_cells, _lib.stream())

    def _enqueue_step(self, e, fresh=False):
        self.backup.copy_(self.q)
        self.cbackup.copy_(self.c)
        st = _lib.stream()
        cs_list = (self.c, self.c1, self.c2)
        cout = (self.c1, self.c2, self.c)
        mxy, mu, mv = self._vel_masks
        tx, ty = self._tiers
        for k in range(3):
            qs = (self.q, self.q1, self.q2)[k]
            self._mix(qs, self.qm, self.Cl2m)
            extrapolated_guess(self.ps, k, e, fresh)
            x = stage_psi_buffer(self.ps, k, e)
            no, ni = self.budgets[k]
            self.batch.enqueue([self.qm[m, 1:-1, 1:-1] for m in range(self.nl)],
                               [x[m, 1:-1, 1:-1] for m in range(self.nl)], rtol=self.rtol,
                               n_outer=no, n_inner=ni, inner_rtol=self.inner_rtol, slot=k)
            self._mix(x, self.psi_l, self.Cm2l)
            _lib.call(f"fvx_geostrophic_faces_{self._csuf}", self.psi_l[0].data_ptr(),
                      _lib.ptr(mxy), _lib.ptr(mu), _lib.ptr(mv), self.u.data_ptr(),
                      self.v.data_ptr(), self.ny + 2, self.nx + 2, self.dx, self.dy, st)
            _lib.call("fvx_qg_stage_batch_launch", self._stage_list(k), st)
            a, b, c = COEF[k]
            _lib.call(f"fvx_advection_stage_{self._csuf}", self.tm, self.tlim,
                      cs_list[k].data_ptr(), self.u.data_ptr(), self.v.data_ptr(),
                      self.c.data_ptr(), cout[k].data_ptr(), _lib.ptr(tx), _lib.ptr(ty),
                      _lib.ptr(self._mh), self.ny + 2, self.nx + 2, self.dx, self.dy, a, b,
                      c * self.dt, st)

    def _restore(self):
        self.q.copy_(self.backup)
        self.c.copy_(self.cbackup)

    def _state(self):
        return [self.q, self.c] + list(self.ps)

    def _build_lists(self):
        for k in range(3):
            self._stage_list(k)


def config5_setup(n: int = 4096, L: float = 1.024e7, seed: int = 0, pv_rms: float = 1e-7):
    """The 3-layer tracer configuration at ``n`` x ``n``: H = [400, 1100,
    2600] m, g' = [9.81, 0.025, 0.0125] m/s^2, the basin-mask generator at
    this size, seeded k^-3 PV per layer and a k^-3-power (amplitude
    |k|^-1.5) tracer.  The PV rms (1e-7 s^-1) keeps the speeds below about
    1.5 m/s over the 200 steps (CFL 0.6 at dt = 900 s, dx = 2.5 km): with the
    barotropic mode's Poisson inversion, the 1e-5 of the 1.5-layer
    configurations gives tens of m/s, beyond the CFL limit on this grid.  The flow evolves freely (no wind forcing,
    ``F0 = 0``): on this 10^4 km beta-plane basin the double-gyre forcing of
    the 1.5-layer model spins up a barotropic western boundary current of
    several m/s within the 200 steps.  Returns the keyword arguments of
    ``MultiLayerQG``."""
    from finitevolx_amd import inputs

    mask = inputs.basin_mask(n, n, 100 + seed)
    # a one-cell coastal wall on the basin's edge: every wet cell then lies
    # where Advection2D writes its tendency ([2:-2] of the padded array), so
    # the flux-form tracer keeps its total exactly
    mask[0, :] = mask[-1, :] = mask[:, 0] = mask[:, -1] = False
    q0 = np.stack([inputs.qg_initial_pv(n, n, rms=pv_rms, seed=200 + 10 * seed + l, mask=mask)
                   for l in range(3)])
    c0 = inputs.spectral_field(n, n, 3.0, 300 + seed) * mask
    return dict(nx=n, ny=n, Lx=L, Ly=L, H=[400.0, 1100.0, 2600.0], g_prime=[9.81, 0.025, 0.0125],
                q0=q0, c0=c0, mask=mask, F0=0.0)
