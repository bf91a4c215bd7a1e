"""Elliptic solvers: masked Helmholtz operator, geometric multigrid and
multigrid-preconditioned (flexible) CG.

Drop-in names for the reference's ``solvers`` package:

* ``masked_laplacian`` (``solvers/iterative.py:144-191``)
* ``build_multigrid_solver`` / ``MultigridSolver`` / ``MultigridLevel``
  (``solvers/multigrid.py:283-1062``), with ``v_cycle``, ``solve_onestep`` and
  ``solve_unrolled``
* ``make_multigrid_preconditioner`` (``solvers/preconditioners.py:207-244``)
* ``solve_cg`` / ``CGInfo`` (``solvers/iterative.py:32-136``)
* ``pv_inversion`` / ``streamfunction_from_vorticity`` (``solvers/elliptic.py:276-542``)
  for ``method="cg"``

The multigrid hierarchy is built natively by ``libfvx_hip.so`` (the same
offline/online split as the reference); ``kind="noflux"`` is the reference
operator (zero face coefficients across land and at the array edge) and
``kind="dirichlet"`` is ``masked_laplacian``'s operator with zero padding at
the array edge.  When ``solve_cg`` receives a multigrid operator and that
solver's preconditioner, the whole PCG loop runs on the device in fused
kernels; other callables take a generic device-side PCG loop.
"""

from __future__ import annotations

import ctypes as C
from typing import NamedTuple

import numpy as np
import torch

from finitevolx_amd import _lib
from finitevolx_amd.mask import Mask2D

_vp, _i, _d, _i64 = _lib._vp, _lib._i, _lib._d, C.c_int64


class _CgInfoC(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer", C.c_int), ("converged", C.c_int),
                ("pad", C.c_int), ("residual_norm", C.c_double), ("rhs_norm", C.c_double)]


_sigs = {
    "fvx_mg_plan_create": [_vp, _vp, _i, _i, _d, _d, _d, _i, _i, _i, _i, _d, _i, _i,
                           C.POINTER(_vp)],
    "fvx_mg_plan_destroy": [_vp],
    "fvx_mg_plan_set_wall": [_vp, _i],
    "fvx_mg_plan_set_chebyshev": [_vp, _d],
    "fvx_mg_plan_info": [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)],
    "fvx_mg_mixed_solve": [_vp, _vp, _i64, _vp, _i64, _d, _d, _i, _i, _d,
                           C.POINTER(_CgInfoC), _vp],
    "fvx_mg_precond_solve": [_vp, _i, _vp, _d, _i, _vp, _i64, _vp, _i64, _d, _d, _i, _i, _d,
                             C.POINTER(_CgInfoC), _vp],
}
PRECONDS = {"multigrid": 0, "learned": 1, "diagonal": 2}
for _s in ("f32", "f64"):
    _sigs.update({
        f"fvx_mg_apply_operator_{_s}": [_vp, _i, _vp, _vp, _vp],
        f"fvx_mg_jacobi_{_s}": [_vp, _i, _vp, _vp, _vp, _i, _vp],
        f"fvx_mg_restrict_{_s}": [_vp, _i, _vp, _vp, _vp],
        f"fvx_mg_prolongate_{_s}": [_vp, _i, _vp, _vp, _vp],
        f"fvx_mg_v_cycle_{_s}": [_vp, _vp, _vp, _vp, _vp],
        f"fvx_mg_v_cycle_level_{_s}": [_vp, _i, _vp, _vp, _vp, _vp],
        f"fvx_mg_pcg_solve_{_s}": [_vp, _vp, _i64, _vp, _i64, _d, _d, _i, _i,
                                  C.POINTER(_CgInfoC), _vp],
    })
_lib.extra_signatures(_sigs)

KINDS = {"noflux": 0, "dirichlet": 1}
PRECISIONS = {"float64": 64, "float32": 32, "mixed": 3264}


class CGInfo(NamedTuple):
    """Convergence diagnostics of ``solve_cg`` (``solvers/iterative.py:32-41``)."""

    iterations: int
    residual_norm: float
    converged: bool


# ---------------------------------------------------------------------------
# masked_laplacian
# ---------------------------------------------------------------------------


def _mask_array(mask, like: torch.Tensor) -> torch.Tensor:
    if isinstance(mask, Mask2D):
        mask = mask.xy_corner_strict
    return _lib.as_field(mask, like.dtype)


def masked_laplacian(psi, mask, dx: float, dy: float, lambda_: float = 0.0):
    """``(lap - lambda) (psi * mask)`` times mask, periodic roll at the array edge."""
    psi = _lib.as_field(psi)
    m = _mask_array(mask, psi)
    if m.shape != psi.shape:
        raise ValueError(f"mask shape {tuple(m.shape)} != psi shape {tuple(psi.shape)}")
    out = torch.empty_like(psi)
    ny, nx = psi.shape
    _lib.call(f"fvx_masked_laplacian_{_lib.suffix(psi)}", psi.data_ptr(), m.data_ptr(),
              out.data_ptr(), ny, nx, float(dx), float(dy), float(lambda_), _lib.stream())
    return out


# ---------------------------------------------------------------------------
# Multigrid
# ---------------------------------------------------------------------------


class MultigridLevel:
    """Handle to one level of a built hierarchy (``multigrid.py:283-322``)."""

    def __init__(self, solver: "MultigridSolver", index: int, shape, dx: float, dy: float):
        self.solver = solver
        self.index = index
        self.shape = tuple(shape)
        self.dx = dx
        self.dy = dy
        self.lambda_ = solver.lambda_

    def __call__(self, u):
        """Apply the level's Helmholtz operator (``_apply_operator``)."""
        return apply_operator(u, self)


def _dtype_of(precision: str) -> torch.dtype:
    return torch.float32 if precision == "float32" else torch.float64


class MultigridSolver:
    """Geometric multigrid V-cycle solver (``multigrid.py:606-902``).

    ``precision``: ``"float64"``, ``"float32"``, or ``"mixed"`` (FP32
    hierarchy for the inner solve plus the FP64 fine operator used for
    iterative refinement; V-cycles then run in FP64 when called directly).
    """

    def __init__(self, handle, ny, nx, dx, dy, lambda_, n_pre, n_post, n_coarse, omega,
                 n_cycles, kind, precision):
        self._h = handle
        self.shape = (ny, nx)
        self.lambda_ = float(lambda_)
        self.n_pre, self.n_post, self.n_coarse = n_pre, n_post, n_coarse
        self.omega, self.n_cycles = omega, n_cycles
        self.kind, self.precision = kind, precision
        nl = C.c_int()
        lc = C.c_int()
        shapes = (C.c_int * 32)()
        _lib.call("fvx_mg_plan_info", self._h, C.byref(nl), shapes, C.byref(lc))
        self.n_levels = nl.value
        self.coarse_level = lc.value
        self.levels = tuple(
            MultigridLevel(self, l, (shapes[2 * l], shapes[2 * l + 1]), dx * 2**l, dy * 2**l)
            for l in range(self.n_levels))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.call("fvx_mg_plan_destroy", h)
            except Exception:
                pass
            self._h = None

    @property
    def dtype(self) -> torch.dtype:
        return _dtype_of(self.precision)

    def _suffix(self, t: torch.Tensor) -> str:
        sfx = _lib.suffix(t)
        if self.precision == "float32" and sfx != "f32":
            raise TypeError("this solver was built for float32 inputs")
        if self.precision == "float64" and sfx != "f64":
            raise TypeError("this solver was built for float64 inputs")
        if self.precision == "mixed" and sfx != "f64":
            raise TypeError("a mixed-precision solver takes float64 inputs")
        return sfx

    def _field(self, x, shape=None):
        t = _lib.as_field(x, self.dtype)
        want = shape or self.shape
        if tuple(t.shape) != tuple(want):
            raise ValueError(f"expected shape {want}, got {tuple(t.shape)}")
        return t

    def v_cycle(self, u, rhs, level_idx: int = 0):
        """One V-cycle from ``u`` (``None`` = zero) on level ``level_idx``
        down to the coarsest (``multigrid.py:808-883``); ``u`` and ``rhs``
        have that level's shape."""
        if not 0 <= level_idx < self.n_levels:
            raise ValueError(f"level_idx must be in [0, {self.n_levels - 1}], got {level_idx}")
        shape = self.levels[level_idx].shape
        rhs = self._field(rhs, shape)
        u = None if u is None else self._field(u, shape)
        out = torch.empty_like(rhs)
        if level_idx == 0:
            _lib.call(f"fvx_mg_v_cycle_{self._suffix(rhs)}", self._h, _lib.ptr(u),
                      rhs.data_ptr(), out.data_ptr(), _lib.stream())
        else:
            _lib.call(f"fvx_mg_v_cycle_level_{self._suffix(rhs)}", self._h, int(level_idx),
                      _lib.ptr(u), rhs.data_ptr(), out.data_ptr(), _lib.stream())
        return out

    def _run_vcycles(self, rhs):
        rhs = self._field(rhs)
        u = None
        for _ in range(self.n_cycles):
            u = self.v_cycle(u, rhs)
        return torch.zeros_like(rhs) if u is None else u

    def __call__(self, rhs):
        """``n_cycles`` V-cycles from zero (``multigrid.py:692-733``)."""
        return self._run_vcycles(rhs)

    # The reference's three solve entry points differ only in how autodiff
    # sees the V-cycles (multigrid.py:692-806); their forward results are
    # the same n_cycles V-cycles from zero.  Gradients through the solve are
    # not provided.
    solve_unrolled = __call__
    solve_onestep = __call__

    # --- fused PCG -------------------------------------------------------
    def pcg(self, rhs, x0=None, rtol=1e-6, atol=1e-6, max_steps=500):
        """Multigrid-preconditioned flexible CG on this hierarchy's fine operator.

        Returns ``(x, info)`` with ``info`` a ``CGInfo``.  For ``precision=
        "mixed"`` the inner solves are FP32 and the residual is certified in FP64.
        """
        rhs_t = _lib.as_field(rhs, self.dtype)
        if tuple(rhs_t.shape) != self.shape:
            raise ValueError(f"expected shape {self.shape}, got {tuple(rhs_t.shape)}")
        x = torch.zeros_like(rhs_t) if x0 is None else _lib.as_field(x0, self.dtype).clone()
        info = self.pcg_into(rhs_t, x, rtol=rtol, atol=atol, max_steps=max_steps,
                             use_x0=x0 is not None)
        return x, info

    def pcg_into(self, rhs: torch.Tensor, x: torch.Tensor, rtol=1e-6, atol=1e-6,
                 max_steps=500, use_x0=True, inner_rtol=1e-4, max_outer=20):
        """In-place solve on (possibly strided 2-D views of) device tensors."""
        ny, nx = self.shape
        for t in (rhs, x):
            if tuple(t.shape) != (ny, nx) or t.stride(1) != 1:
                raise ValueError("rhs and x must be (ny, nx) views with unit column stride")
        ci = _CgInfoC()
        if self.precision == "mixed":
            _lib.call("fvx_mg_mixed_solve", self._h, rhs.data_ptr(), rhs.stride(0), x.data_ptr(),
                      x.stride(0), float(rtol), float(atol), int(max_outer), int(max_steps),
                      float(inner_rtol), C.byref(ci), _lib.stream())
        else:
            _lib.call(f"fvx_mg_pcg_solve_{self._suffix(rhs)}", self._h, rhs.data_ptr(),
                      rhs.stride(0), x.data_ptr(), x.stride(0), float(rtol), float(atol),
                      int(max_steps), 1 if use_x0 else 0, C.byref(ci), _lib.stream())
        self.last_info = ci
        return CGInfo(iterations=int(ci.iterations), residual_norm=float(ci.residual_norm),
                      converged=bool(ci.converged))

    def precond_solve(self, rhs, x0=None, precond: str = "multigrid", net=None,
                      out_scale: float = 0.0, sweeps: int = 2, rtol=1e-8, atol=0.0,
                      max_outer=20, max_inner=200, inner_rtol=1e-4):
        """Mixed-precision PCG with the inner preconditioner chosen by name:
        ``"multigrid"`` (this hierarchy's V-cycle), ``"learned"`` (a trained
        ``learned.UNet``; with ``sweeps`` > 0 it corrects the smooth error
        between that many damped-Jacobi pre- and post-sweeps), or
        ``"diagonal"`` (Jacobi).  Needs ``precision="mixed"``.  Returns
        ``(x, info)``; ``self.last_info.pad`` counts inner solves that hit
        ``max_inner`` and were redone with the V-cycle."""
        if self.precision != "mixed":
            raise ValueError("precond_solve needs a solver built with precision='mixed'")
        if precond not in PRECONDS:
            raise ValueError(f"precond must be one of {list(PRECONDS)}, got {precond!r}")
        if precond == "learned" and net is None:
            raise ValueError("precond='learned' needs a network")
        rhs_t = _lib.as_field(rhs, torch.float64)
        if tuple(rhs_t.shape) != self.shape:
            raise ValueError(f"expected shape {self.shape}, got {tuple(rhs_t.shape)}")
        x = torch.zeros_like(rhs_t) if x0 is None else _lib.as_field(x0, torch.float64).clone()
        ci = _CgInfoC()
        _lib.call("fvx_mg_precond_solve", self._h, PRECONDS[precond],
                  None if net is None else net._h, float(out_scale), int(sweeps), rhs_t.data_ptr(),
                  rhs_t.stride(0), x.data_ptr(), x.stride(0), float(rtol), float(atol),
                  int(max_outer), int(max_inner), float(inner_rtol), C.byref(ci), _lib.stream())
        self.last_info = ci
        return x, CGInfo(iterations=int(ci.iterations), residual_norm=float(ci.residual_norm),
                         converged=bool(ci.converged))


def apply_operator(u, level: MultigridLevel):
    """``_apply_operator`` (``multigrid.py:327-405``) at one level."""
    s = level.solver
    u = s._field(u, level.shape)
    out = torch.empty_like(u)
    _lib.call(f"fvx_mg_apply_operator_{s._suffix(u)}", s._h, level.index, u.data_ptr(),
              out.data_ptr(), _lib.stream())
    return out


def weighted_jacobi(u, rhs, level: MultigridLevel, omega: float, n_iters: int):
    """``_weighted_jacobi`` (``multigrid.py:408-461``) with the solver's omega."""
    s = level.solver
    if abs(float(omega) - s.omega) > 0.0:
        raise ValueError(f"omega {omega} differs from the solver's omega {s.omega}")
    rhs = s._field(rhs, level.shape)
    u = None if u is None else s._field(u, level.shape)
    out = torch.empty_like(rhs)
    _lib.call(f"fvx_mg_jacobi_{s._suffix(rhs)}", s._h, level.index, _lib.ptr(u),
              rhs.data_ptr(), out.data_ptr(), int(n_iters), _lib.stream())
    return out


def restrict(v, level: MultigridLevel):
    """``_restrict`` (``multigrid.py:464-507``) from ``level`` to the next coarser."""
    s = level.solver
    v = s._field(v, level.shape)
    cs = s.levels[level.index + 1].shape
    out = torch.empty(cs, dtype=v.dtype, device=v.device)
    _lib.call(f"fvx_mg_restrict_{s._suffix(v)}", s._h, level.index, v.data_ptr(),
              out.data_ptr(), _lib.stream())
    return out


def prolongate(vc, level: MultigridLevel):
    """``_prolongate`` (``multigrid.py:510-603``) onto ``level`` from the next coarser."""
    s = level.solver
    vc = s._field(vc, s.levels[level.index + 1].shape)
    out = torch.empty(level.shape, dtype=vc.dtype, device=vc.device)
    _lib.call(f"fvx_mg_prolongate_{s._suffix(vc)}", s._h, level.index, vc.data_ptr(),
              out.data_ptr(), _lib.stream())
    return out


def build_multigrid_solver(mask, dx: float, dy: float, lambda_: float = 0.0, coeff=None,
                           n_levels: int | None = None, n_pre: int = 6, n_post: int = 6,
                           n_coarse: int = 50, omega: float = 0.95, n_cycles: int = 5, *,
                           kind: str = "noflux", precision: str = "float64",
                           wall: bool = False, chebyshev: float | None = None) -> MultigridSolver:
    """Offline hierarchy build (``multigrid.py:905-1062``).

    ``mask`` may be a ``Mask2D`` (its ``xy_corner_strict`` is used, as in the
    reference) or any 0/1 array.  Non-divisible grids raise ``ValueError``.
    ``wall`` (mixed-precision Dirichlet plans whose levels are all tiled; no
    reference counterpart): the coarse-level wall correction of the V-cycle.
    ``chebyshev`` (any plan with 1-3 sweeps per leg; no reference
    counterpart): the lower end ``lo`` of the smoothing interval [lo, 2]
    whose Chebyshev weights replace the uniform ``omega`` in the V-cycle's
    sweeps.
    """
    if kind not in KINDS:
        raise ValueError(f"kind must be one of {list(KINDS)}, got {kind!r}")
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {list(PRECISIONS)}, got {precision!r}")
    if isinstance(mask, Mask2D):
        m = np.asarray(mask.xy_corner_strict, dtype=np.float64)
    elif isinstance(mask, torch.Tensor):
        m = mask.detach().cpu().numpy().astype(np.float64)
    else:
        m = np.asarray(mask, dtype=np.float64)
    if m.ndim != 2:
        raise ValueError(f"mask must be 2-D, got shape {m.shape}")
    m = np.ascontiguousarray(m)
    c = None
    if coeff is not None:
        c = coeff.detach().cpu().numpy() if isinstance(coeff, torch.Tensor) else np.asarray(coeff)
        c = np.ascontiguousarray(c, dtype=np.float64)
        if c.shape != m.shape:
            raise ValueError(f"coeff shape {c.shape} != mask shape {m.shape}")
    _lib.require_gpu()
    ny, nx = m.shape
    h = _vp()
    _lib.call("fvx_mg_plan_create", m.ctypes.data, None if c is None else c.ctypes.data, ny, nx,
              float(dx), float(dy), float(lambda_), int(n_levels or 0), int(n_pre), int(n_post),
              int(n_coarse), float(omega), KINDS[kind], PRECISIONS[precision], C.byref(h))
    solver = MultigridSolver(h, ny, nx, dx, dy, lambda_, n_pre, n_post, n_coarse, omega, n_cycles,
                             kind, precision)
    if wall:
        _lib.call("fvx_mg_plan_set_wall", h, 1)
    if chebyshev is not None:
        _lib.call("fvx_mg_plan_set_chebyshev", h, float(chebyshev))
    return solver


# ---------------------------------------------------------------------------
# Preconditioner and CG
# ---------------------------------------------------------------------------


class MultigridPreconditioner:
    """``M^{-1} r = V(0, r)`` (``preconditioners.py:207-244``)."""

    def __init__(self, solver: MultigridSolver):
        self.solver = solver

    def __call__(self, r):
        return self.solver.v_cycle(None, r)


def make_multigrid_preconditioner(mg_solver: MultigridSolver) -> MultigridPreconditioner:
    return MultigridPreconditioner(mg_solver)


class MultigridOperator:
    """The fine-level operator of a solver as a matvec callable."""

    def __init__(self, solver: MultigridSolver):
        self.solver = solver

    def __call__(self, u):
        return apply_operator(u, self.solver.levels[0])


def _operator_solver(matvec):
    if isinstance(matvec, MultigridOperator):
        return matvec.solver
    if isinstance(matvec, MultigridLevel) and matvec.index == 0:
        return matvec.solver
    return None


def solve_cg(matvec, rhs, x0=None, preconditioner=None, rtol: float = 1e-6, atol: float = 1e-6,
             max_steps: int | None = 500):
    """Preconditioned CG for the negative-definite ``A`` (``iterative.py:48-136``).

    Stopping rule: ``||b - A x||_2 <= max(rtol ||b||_2, atol)`` (the
    reference's lineax 0.1.0 rule is not restated).  With a multigrid
    operator and that solver's preconditioner, the fused device path runs.
    Returns ``(x, CGInfo)``.
    """
    max_steps = 10**6 if max_steps is None else int(max_steps)
    mg = _operator_solver(matvec)
    if (mg is not None and isinstance(preconditioner, MultigridPreconditioner)
            and preconditioner.solver is mg):
        return mg.pcg(rhs, x0=x0, rtol=rtol, atol=atol, max_steps=max_steps)
    return _generic_pcg(matvec, rhs, x0, preconditioner, rtol, atol, max_steps)


def _generic_pcg(matvec, rhs, x0, precond, rtol, atol, max_steps, check_every=8):
    """Flexible PCG for arbitrary device callables (vector algebra in torch).

    The stopping test runs on the device every iteration; the host reads it
    every ``check_every`` iterations.  Iterations after the test first holds
    leave ``x`` unchanged, so the result and the iteration count are those
    of a host test after every iteration."""
    b = _lib.as_field(rhs)
    x = torch.zeros_like(b) if x0 is None else _lib.as_field(x0, b.dtype).clone()
    M = precond if precond is not None else (lambda r: r)
    r = b - matvec(x)
    bn = float(torch.linalg.vector_norm(b))
    tol = max(rtol * bn, atol)
    rn = float(torch.linalg.vector_norm(r))
    if rn <= tol:
        return x, CGInfo(0, rn, True)
    z = _lib.as_field(M(r), b.dtype)
    p = z.clone()
    rz = torch.sum(r * z)
    live = torch.ones((), dtype=torch.bool, device=b.device)
    its = torch.zeros((), dtype=torch.int64, device=b.device)
    zero = torch.zeros((), dtype=b.dtype, device=b.device)
    for k in range(1, max_steps + 1):
        q = matvec(p)
        alpha = torch.where(live, rz / torch.sum(p * q), zero)
        x = x + alpha * p
        r_old = r
        r = r - alpha * q
        its += live
        live = live & (torch.linalg.vector_norm(r) > tol)
        if (k % check_every == 0 or k == max_steps) and not bool(live):
            break
        z = _lib.as_field(M(r), b.dtype)
        rz_new = torch.sum(r * z)
        beta = torch.where(live, torch.sum(z * (r - r_old)) / rz, zero)
        rz = torch.where(live, rz_new, rz)
        p = z + beta * p
    res = float(torch.linalg.vector_norm(b - matvec(x)))
    return x, CGInfo(int(its), res, not bool(live))


# ---------------------------------------------------------------------------
# Spectral rectangular solvers (solvers/spectral.py:14-128 re-exports of
# spectraldiffx v0.0.10): each inverts one 5-point operator exactly.
# ---------------------------------------------------------------------------

_lib.extra_signatures({
    "fvx_spectral_plan_create": [_i, _i, _i, _i, C.POINTER(_vp)],
    "fvx_spectral_plan_destroy": [_vp],
    "fvx_spectral_helmholtz_f64": [_vp, _vp, _i64, _vp, _i64, _d, _d, _d, _vp],
    "fvx_spectral_helmholtz_f32": [_vp, _vp, _i64, _vp, _i64, _d, _d, _d, _vp],
})

_BC = {"dst": 0, "dct": 1, "fft": 2}


class _SpecPlan:
    def __init__(self, ny, nx, bc, bits):
        h = _vp()
        _lib.call("fvx_spectral_plan_create", ny, nx, _BC[bc], bits, C.byref(h))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            try:
                _lib.call("fvx_spectral_plan_destroy", self.h)
            except Exception:
                pass
            self.h = None


_SPEC_PLANS: dict = {}


def _spectral(rhs, dx, dy, lambda_, bc):
    b = _lib.as_field(rhs)
    if b.dim() < 2:
        raise ValueError("spectral solves need at least 2-D input")
    ny, nx = b.shape[-2:]
    bits = 64 if b.dtype == torch.float64 else 32
    key = (ny, nx, bc, bits)
    plan = _SPEC_PLANS.get(key)
    if plan is None:
        if len(_SPEC_PLANS) > 8:
            _SPEC_PLANS.clear()
        plan = _SPEC_PLANS[key] = _SpecPlan(ny, nx, bc, bits)
    flat = b.reshape(-1, ny, nx)
    out = torch.empty_like(flat)
    for k in range(flat.shape[0]):
        _lib.call(f"fvx_spectral_helmholtz_{_lib.suffix(b)}", plan.h, flat[k].data_ptr(), nx,
                  out[k].data_ptr(), nx, float(dx), float(dy), float(lambda_), _lib.stream())
    return out.reshape(b.shape)


def solve_helmholtz_dst(rhs, dx, dy, lambda_=0.0):
    """Dirichlet ``(lap - lambda) psi = rhs`` on an interior-only array (DST-I)."""
    return _spectral(rhs, dx, dy, lambda_, "dst")


def solve_helmholtz_dct(rhs, dx, dy, lambda_=0.0):
    """Neumann (no-flux, cell-centred) Helmholtz solve (DCT-II); the mean
    mode is set to zero when ``lambda_ == 0``."""
    return _spectral(rhs, dx, dy, lambda_, "dct")


def solve_helmholtz_fft(rhs, dx, dy, lambda_=0.0):
    """Periodic Helmholtz solve (2-D real FFT); zero mean mode when singular."""
    return _spectral(rhs, dx, dy, lambda_, "fft")


def solve_poisson_dst(rhs, dx, dy):
    return solve_helmholtz_dst(rhs, dx, dy, 0.0)


def solve_poisson_dct(rhs, dx, dy):
    return solve_helmholtz_dct(rhs, dx, dy, 0.0)


def solve_poisson_fft(rhs, dx, dy):
    return solve_helmholtz_fft(rhs, dx, dy, 0.0)


def dst1_eigenvalues(n: int, d: float) -> np.ndarray:
    """Eigenvalues of the 1-D Dirichlet second difference (DST-I basis)."""
    k = np.arange(1, n + 1)
    return -4.0 / d**2 * np.sin(np.pi * k / (2.0 * (n + 1))) ** 2


_HELMHOLTZ_DISPATCH = {"dst": solve_helmholtz_dst, "dct": solve_helmholtz_dct,
                       "fft": solve_helmholtz_fft}


# ---------------------------------------------------------------------------
# Wrappers (elliptic.py) for method="cg"
# ---------------------------------------------------------------------------


def _interior_view(mask_arr: np.ndarray):
    """True when the outer ring is dry, so the Dirichlet solve fits the interior."""
    return (mask_arr.shape[0] > 2 and mask_arr.shape[1] > 2 and not mask_arr[0, :].any()
            and not mask_arr[-1, :].any() and not mask_arr[:, 0].any()
            and not mask_arr[:, -1].any())


def _resolve_mask(mask):
    if mask is None:
        return None
    if isinstance(mask, Mask2D):
        return np.asarray(mask.xy_corner_strict, dtype=np.float64)
    if isinstance(mask, torch.Tensor):
        return mask.detach().cpu().numpy().astype(np.float64)
    return np.asarray(mask, dtype=np.float64)


_SOLVER_CACHE: dict = {}


def _dirichlet_solver(m: np.ndarray, dx, dy, lam, precision):
    key = (m.shape, m.tobytes(), float(dx), float(dy), float(lam), precision)
    s = _SOLVER_CACHE.get(key)
    if s is None:
        s = build_multigrid_solver(m, dx, dy, lambda_=lam, n_pre=2, n_post=2, n_coarse=20,
                                   kind="dirichlet", precision=precision)
        if len(_SOLVER_CACHE) > 16:
            _SOLVER_CACHE.clear()
        _SOLVER_CACHE[key] = s
    return s


def _solve_cg_method(rhs, dx, dy, lam, mask, preconditioner, rtol=1e-10, atol=0.0):
    m = _resolve_mask(mask)
    if m is None:
        raise ValueError("method='cg' requires a mask (array or Mask2D)")
    b = _lib.as_field(rhs)
    if preconditioner is not None:
        mt = _lib.as_field(m, b.dtype)
        x, _ = _generic_pcg(lambda x: masked_laplacian(x, mt, dx, dy, lam), b * mt, None,
                            preconditioner, rtol, atol, 10**5)
        return x * mt
    prec = "float64" if b.dtype == torch.float64 else "float32"
    out = torch.zeros_like(b)
    if _interior_view(m):
        inner = m[1:-1, 1:-1]
        s = _dirichlet_solver(np.ascontiguousarray(inner), dx, dy, lam, prec)
        s.pcg_into(b[1:-1, 1:-1], out[1:-1, 1:-1], rtol=rtol, atol=atol, max_steps=1000,
                   use_x0=False)
    else:
        s = _dirichlet_solver(m, dx, dy, lam, prec)
        s.pcg_into(b, out, rtol=rtol, atol=atol, max_steps=1000, use_x0=False)
    return out * _lib.as_field(m, b.dtype)


def _spectral_dispatch(rhs, dx, dy, lam, bc):
    fn = _HELMHOLTZ_DISPATCH.get(bc)
    if fn is None:
        raise ValueError(f"bc must be 'fft', 'dst', or 'dct'; got {bc!r}")
    return fn(rhs, dx, dy, lam)


def streamfunction_from_vorticity(zeta, dx, dy, bc="dst", lambda_=0.0, method="spectral",
                                  mask=None, capacitance_solver=None, preconditioner=None):
    """``(lap - lambda) psi = zeta`` (``elliptic.py:276-339``)."""
    if method == "cg":
        return _solve_cg_method(zeta, dx, dy, lambda_, mask, preconditioner)
    if method == "spectral":
        return _spectral_dispatch(zeta, dx, dy, lambda_, bc)
    if method == "capacitance":
        raise NotImplementedError("method='capacitance' is outside this build's hot path")
    raise ValueError(f"method must be 'spectral', 'cg', or 'capacitance'; got {method!r}")


def pv_inversion(pv, dx, dy, lambda_, bc="dst", method="spectral", mask=None,
                 capacitance_solver=None, preconditioner=None):
    """QG PV inversion ``(lap - lambda) psi = q`` (``elliptic.py:390-542``).

    Scalar ``lambda_`` solves every leading slice; an array ``lambda_`` of
    length ``nl`` pairs layer ``k`` of ``pv[..., k, :, :]`` with ``lambda_[k]``.
    """
    if method == "capacitance":
        raise NotImplementedError("method='capacitance' is outside this build's hot path")
    if method not in ("cg", "spectral"):
        raise ValueError(f"method must be 'spectral', 'cg', or 'capacitance'; got {method!r}")
    pv = _lib.as_field(pv)
    lam = np.asarray(lambda_, dtype=np.float64)

    def one(f, lk):
        if method == "spectral":
            return _spectral_dispatch(f, dx, dy, lk, bc)
        return _solve_cg_method(f, dx, dy, lk, mask, preconditioner)

    if lam.ndim == 0:
        flat = pv.reshape(-1, *pv.shape[-2:])
        out = torch.stack([one(f, float(lam)) for f in flat])
        return out.reshape(pv.shape)
    if pv.dim() < 3:
        raise ValueError(f"pv must have at least 3 dims when lambda_ is an array, got shape "
                         f"{tuple(pv.shape)}")
    nl = lam.shape[0]
    if pv.shape[-3] != nl:
        raise ValueError(f"pv.shape[-3]={pv.shape[-3]} does not match lambda_ length {nl}")
    flat = pv.reshape(-1, nl, *pv.shape[-2:])
    out = torch.stack([torch.stack([one(f[k], float(lam[k])) for k in range(nl)])
                       for f in flat])
    return out.reshape(pv.shape)


# ---------------------------------------------------------------------------
# Work-list batches of independent mixed-precision solves
# ---------------------------------------------------------------------------

_lib.extra_signatures({
    "fvx_mg_batch_create": [C.POINTER(_vp), _i, C.POINTER(_vp)],
    "fvx_mg_batch_destroy": [_vp],
    "fvx_mg_batch_mixed_enqueue": [_vp, C.POINTER(_vp), C.POINTER(_i64), C.POINTER(_vp),
                                   C.POINTER(_i64), _d, _d, _i, C.POINTER(_i), _d, _i, _vp],
    "fvx_mg_batch_status": [_vp, _i, _i, C.POINTER(_CgInfoC), C.POINTER(_i), _vp],
})


class MultigridBatch:
    """Independent mixed-precision multigrid-PCG solves of several
    ``MultigridSolver`` plans (``precision="mixed"``; any sizes, masks and
    lambdas, the same operator kind and sweep counts) run together: each
    operator of the solve is one work-list launch over all members, and each
    member stops on its own FP64-certified convergence.  Batches the
    reference's per-layer ``vmap`` solves (``solvers/elliptic.py:532-541``)
    and ensembles of independent models."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        if not self.solvers:
            raise ValueError("an empty batch")
        for s in self.solvers:
            if s.precision != "mixed":
                raise ValueError("batched solves need precision='mixed' plans")
        arr = (_vp * len(self.solvers))(*[s._h for s in self.solvers])
        h = _vp()
        _lib.call("fvx_mg_batch_create", arr, len(self.solvers), C.byref(h))
        self._h = h
        self.n = len(self.solvers)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.call("fvx_mg_batch_destroy", h)
            except Exception:
                pass
            self._h = None

    def enqueue(self, rhs, x, rtol=1e-8, atol=0.0, n_outer=2, n_inner=(8, 8),
                inner_rtol=1e-4, slot=0):
        """Enqueue the solves ``A_m x_m = rhs_m`` (FP64 (ny, nx) views with unit
        column stride; ``x`` is the warm start, updated in place) on ``slot``
        with ``n_outer`` refinement steps of ``n_inner[o]`` inner iterations
        (device-guarded: converged members skip the rest).  No host sync."""
        if len(rhs) != self.n or len(x) != self.n:
            raise ValueError(f"need {self.n} right-hand sides and solutions")
        for s, b, xx in zip(self.solvers, rhs, x):
            for t in (b, xx):
                if tuple(t.shape) != s.shape or t.dtype != torch.float64 or t.stride(1) != 1:
                    raise ValueError("rhs / x must be float64 (ny, nx) views of each member")
        ni = (_i * len(n_inner))(*[int(k) for k in n_inner])
        if len(n_inner) < n_outer:
            raise ValueError("n_inner needs one entry per refinement step")
        _lib.call("fvx_mg_batch_mixed_enqueue", self._h,
                  (_vp * self.n)(*[b.data_ptr() for b in rhs]),
                  (_i64 * self.n)(*[b.stride(0) for b in rhs]),
                  (_vp * self.n)(*[xx.data_ptr() for xx in x]),
                  (_i64 * self.n)(*[xx.stride(0) for xx in x]), float(rtol), float(atol),
                  int(n_outer), ni, float(inner_rtol), int(slot), _lib.stream())

    def status(self, slot=0, n_outer=2):
        """Per member: ``(CGInfo, refinement steps, inner counts per step)``
        of the last solve on ``slot`` (synchronises)."""
        infos = (_CgInfoC * self.n)()
        counts = (_i * (4 * self.n))()
        _lib.call("fvx_mg_batch_status", self._h, int(slot), int(n_outer), infos, counts,
                  _lib.stream())
        out = []
        for m in range(self.n):
            ci = infos[m]
            out.append((CGInfo(iterations=int(ci.iterations), residual_norm=float(ci.residual_norm),
                               converged=bool(ci.converged)), int(ci.outer),
                        [counts[4 * m + o] for o in range(4)]))
        return out
