"""1.5-layer quasi-geostrophic model on the GPU (the fused hot path).

Build-defined composition of reference operators (no single reference
function; nearest: ``scripts/qg_1p5_layer.py:420-497``)::

    dq/dt = -J(psi, q + beta (y - y0)) + F(y) - nu4 D(D(q))
    (lap - 1/Ld^2) psi = q      (psi = 0 on land and beyond the walls)

stepped by ``rk3_ssp_step``.  Each RK stage is one Helmholtz solve
(multigrid-preconditioned flexible CG; FP64, or FP32 inner solves with FP64
iterative refinement) and ONE fused HIP kernel for the tendency plus the stage
combination.  The whole step loop runs natively (``fvx_qg_step``).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from finitevolx_amd import _lib

_vp, _i, _d = _lib._vp, _lib._i, _lib._d


class _QGInfoC(C.Structure):
    _fields_ = [("solves", C.c_int), ("iterations", C.c_int), ("max_iterations", C.c_int),
                ("outer", C.c_int), ("converged", C.c_int), ("retries", C.c_int),
                ("max_rel_residual", C.c_double)]


_lib.extra_signatures({
    "fvx_qg_plan_create": [_i, _i, _d, _d, _d, _d, _d, _d, _d, _vp, _i, _i, _i, _i, _i, _d, _d,
                           _d, _i, _i, C.POINTER(_vp)],
    "fvx_qg_plan_destroy": [_vp],
    "fvx_qg_step": [_vp, _vp, _vp, _i, C.POINTER(_QGInfoC), _vp],
    "fvx_qg_step_graph": [_vp, _vp, _vp, _i, C.POINTER(_QGInfoC), _vp],
    "fvx_qg_set_preconditioner": [_vp, _vp, _d, _i],
    "fvx_qg_fallbacks": [_vp],
    "fvx_qg_budgets": [_vp, C.POINTER(C.c_int)],
    "fvx_qg_set_warm_start": [_vp, _i],
    "fvx_qg_set_wall": [_vp, _i, C.POINTER(_i)],
    "fvx_qg_set_chebyshev": [_vp, _d, C.POINTER(_i)],
    "fvx_qg_stage_f64": [_vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _d, _d, _d, _d, _d, _vp],
    "fvx_qg_stage_f32": [_vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _d, _d, _d, _d, _d, _vp],
    "fvx_qg_stage_f64t32": [_vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _d, _d, _d, _d, _d,
                            _vp],
})

# precision name -> (state code, solve mode); state code 6432 = FP64 state
# with the tendency computed in FP32
MODES = {"float64": (64, 64), "mixed": (6432, 3264), "mixed32": (32, 3264),
         "mixed64": (64, 3264)}


class QGInfo(dict):
    """Per-call solver statistics (solves, iterations, outer sweeps, residual)."""


class QGModel:
    """Fused GPU QG model on a closed ``ny x nx`` basin (interior cells).

    ``precision``: ``"float64"`` (FP64 state, stencils and MG-PCG);
    ``"mixed"`` (FP64 state, tendency stencils in FP32, FP32 MG-PCG inner
    solves with FP64 refinement of psi); ``"mixed32"`` (as ``"mixed"`` with
    the state stored in FP32, whose rounding each stage drifts the PV by
    ~1e-6 relative over 100 steps); ``"mixed64"`` (as ``"mixed"`` with the
    tendency in FP64).
    ``mask``: optional interior wet mask ``(ny, nx)``.
    """

    def __init__(self, nx: int, ny: int, Lx: float, Ly: float, *, beta: float = 1.754e-11,
                 Ld: float = 4e4, nu4: float = 1e9, F0: float = 2e-12, dt: float = 1200.0,
                 mask=None, precision: str = "float64", n_pre: int = 3, n_post: int = 3,
                 n_coarse: int = 3, omega: float = 0.85, rtol: float = 1e-8,
                 inner_rtol: float = 1e-4, max_iters: int = 100, max_outer: int = 10):
        if precision not in MODES:
            raise ValueError(f"precision must be one of {list(MODES)}")
        _lib.require_gpu()
        self.nx, self.ny, self.Lx, self.Ly = nx, ny, float(Lx), float(Ly)
        self.dx, self.dy = Lx / nx, Ly / ny
        self.beta, self.Ld, self.nu4, self.F0, self.dt = beta, Ld, nu4, F0, dt
        self.lambda_ = 1.0 / Ld**2
        self.precision = precision
        self.state_code, solve = MODES[precision]
        self.state_bits = 32 if self.state_code == 32 else 64
        self.dtype = torch.float64 if self.state_bits == 64 else torch.float32
        m = None
        self.mask = None
        if mask is not None:
            m = np.ascontiguousarray(np.asarray(mask, dtype=np.float64))
            if m.shape != (ny, nx):
                raise ValueError(f"mask must be the interior shape {(ny, nx)}, got {m.shape}")
            self.mask = m > 0
        h = _vp()
        _lib.call("fvx_qg_plan_create", ny, nx, float(Lx), float(Ly), float(beta), float(Ld),
                  float(nu4), float(F0), float(dt), None if m is None else m.ctypes.data,
                  self.state_code, solve, n_pre, n_post, n_coarse, float(omega), float(rtol),
                  float(inner_rtol), int(max_iters), int(max_outer), C.byref(h))
        self._h = h
        full = np.zeros((ny + 2, nx + 2), dtype=np.uint8)
        full[1:-1, 1:-1] = 1 if m is None else (m > 0)
        self._mask_full = None if m is None else torch.from_numpy(full).to(_lib.require_gpu())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.call("fvx_qg_plan_destroy", h)
            except Exception:
                pass
            self._h = None

    @property
    def shape(self):
        return (self.ny + 2, self.nx + 2)

    def set_preconditioner(self, net=None, out_scale: float = 0.0, sweeps: int = 2):
        """Precondition the inner FP32 solves (mixed precision only) with the
        learned two-grid cycle: ``sweeps`` damped-Jacobi pre/post sweeps and a
        trained ``learned.UNet`` (sized for the half-resolution grid) as the
        coarse solve; ``None`` restores the multigrid V-cycle."""
        if net is not None and (net.H < self.ny // 2 or net.W < self.nx // 2):
            raise ValueError("the network plan is smaller than the half-resolution grid")
        self._net = net  # keep the plan alive while the model uses it
        _lib.call("fvx_qg_set_preconditioner", self._h, None if net is None else net._h,
                  float(out_scale), int(sweeps))

    def set_wall_correction(self, on: bool = True) -> bool:
        """Coarse-level wall correction of the inner FP32 V-cycles (mixed
        precision; no reference counterpart): on level l each dry or outside
        neighbour takes the ghost value g_l u, g_l = (1 - 2^l) / (1 + 2^l), so
        every level keeps the fine operator's effective wall.  Returns whether
        it is on (hierarchies with a coarse level below 64 x 64 cells, and the
        FP64 path, cannot take it).  The FP64 certification is unchanged."""
        en = _i(0)
        _lib.call("fvx_qg_set_wall", self._h, int(bool(on)), C.byref(en))
        return bool(en.value)

    def set_chebyshev(self, lo: float | None = 0.1) -> bool:
        """Chebyshev weights for the V-cycles' damped-Jacobi sweeps (no
        reference counterpart): sweep k of an m-sweep leg takes the
        reciprocal of the k-th root of the degree-m Chebyshev polynomial on
        [lo, 2].  ``None`` restores the uniform weight.  Returns whether it
        is on (plans with more than 3 sweeps per leg cannot take it).  The
        FP64 certification is unchanged.  A mixed plan starts with it on
        where it can (lo = 0.25 with a land mask, 0.4 without); an FP64 plan
        starts with the uniform weight (the reference's smoother)."""
        en = _i(0)
        _lib.call("fvx_qg_set_chebyshev", self._h, 0.0 if lo is None else float(lo), C.byref(en))
        return bool(en.value)

    def set_warm_start(self, mode: str = "extrapolate"):
        """Warm start of each stage's Helmholtz solve: ``"extrapolate"``
        (default; linear-in-time extrapolation from the previous stages'
        streamfunctions, exact for psi linear in time) or ``"previous"`` (the
        previous stage's streamfunction)."""
        modes = {"previous": 0, "extrapolate": 1}
        if mode not in modes:
            raise ValueError(f"warm-start mode must be one of {list(modes)}")
        _lib.call("fvx_qg_set_warm_start", self._h, modes[mode])

    @property
    def budgets(self) -> dict:
        """Iteration budgets of the captured step graphs and recaptures so far."""
        out = (C.c_int * 4)()
        _lib.call("fvx_qg_budgets", self._h, out)
        return dict(iters=out[0], inner=out[1], outer=out[2], recaptures=out[3])

    @property
    def fallbacks(self) -> int:
        """Times the learned preconditioner gave way to multigrid-PCG."""
        return _lib.call("fvx_qg_fallbacks", self._h)

    def new_state(self, q_interior):
        """Full ``(ny+2, nx+2)`` state with a zero ghost ring (and zero land)."""
        q = torch.zeros(self.shape, dtype=self.dtype, device=_lib.require_gpu())
        qi = _lib.as_field(q_interior, self.dtype)
        if self.mask is not None:
            qi = qi * torch.from_numpy(self.mask.astype(np.float64)).to(qi.device, qi.dtype)
        q[1:-1, 1:-1] = qi
        return q

    def step(self, q: torch.Tensor, psi: torch.Tensor | None = None, nsteps: int = 1,
             graph: bool = True):
        """Advance ``q`` in place by ``nsteps`` RK3 steps; returns ``(q, psi, info)``.

        ``psi`` (FP64, full shape) is used as the warm start and updated.
        ``graph=True`` replays one captured HIP graph per step (solver budgets
        calibrated on the first call, every solve certified by its FP64
        residual flag); ``graph=False`` drives each solve from the host.
        """
        if q.shape != self.shape or q.dtype != self.dtype or not q.is_cuda or not q.is_contiguous():
            raise ValueError(f"q must be a contiguous {self.dtype} CUDA tensor of shape {self.shape}")
        if psi is None:
            psi = torch.zeros(self.shape, dtype=torch.float64, device=q.device)
        if psi.shape != self.shape or psi.dtype != torch.float64 or not psi.is_contiguous():
            raise ValueError("psi must be a contiguous float64 tensor of the full shape")
        ci = _QGInfoC()
        _lib.call("fvx_qg_step_graph" if graph else "fvx_qg_step", self._h, q.data_ptr(),
                  psi.data_ptr(), int(nsteps), C.byref(ci), _lib.stream())
        info = QGInfo(solves=ci.solves, iterations=ci.iterations,
                      max_iterations=ci.max_iterations, outer=ci.outer,
                      converged=bool(ci.converged), retries=ci.retries,
                      max_rel_residual=ci.max_rel_residual)
        return q, psi, info

    def stage(self, q_s, psi, q_n, a: float, b: float, c: float):
        """One fused tendency + stage combination ``a q_n + b q_s + c dt rhs`` (given psi)."""
        q_s = _lib.as_field(q_s, self.dtype)
        q_n = _lib.as_field(q_n, self.dtype)
        psi = _lib.as_field(psi, torch.float64)
        out = torch.empty_like(q_s)
        Ny, Nx = self.shape
        suf = "f64t32" if self.state_code == 6432 else _lib.suffix(q_s)
        _lib.call(f"fvx_qg_stage_{suf}", q_s.data_ptr(), psi.data_ptr(),
                  q_n.data_ptr(), out.data_ptr(), _lib.ptr(self._mask_full), Ny, Nx, self.dx,
                  self.dy, self.beta, self.Ly, self.F0, self.nu4, float(a), float(b),
                  float(c) * self.dt, _lib.stream())
        return out
