"""Flux-form advection (drop-in for ``Advection2D``, ``advection/advection.py:283-420``).

Every method runs as one fused HIP kernel: LDS-staged tracer tile,
one upwind-selected reconstruction per face, flux divergence, mask and the
zero 2-cell output ring in the same pass.  With a mask, the WENO3/WENO5/
WENO-Z5/TVD methods dispatch per face on the upwind cell's stencil tier
(``advection/flux.py:100-253``); other methods run unmasked and the result
is multiplied by ``mask.h`` (``advection.py:363-419``).
"""

from __future__ import annotations

import torch

from finitevolx_amd import _lib
from finitevolx_amd.grid import CartesianGrid2D
from finitevolx_amd.mask import Mask2D

# Method and limiter codes (FVX_ADV_* / FVX_LIM_* in include/fvx.h)
# ("wenoz3", code 5, exists only as the 4-point tier of "wenoz5", as in the
# reference's method table, advection.py:363-401.)
_METHODS = {"naive": 0, "upwind1": 1, "upwind2": 2, "upwind3": 3, "weno3": 4,
            "weno5": 6, "wenoz5": 7}
_LIMITERS = {"minmod": 0, "van_leer": 1, "superbee": 2, "mc": 3}
_MASK_DISPATCHABLE = frozenset({"weno3", "weno5", "wenoz5"}) | frozenset(_LIMITERS)
_UNSUPPORTED = frozenset({"weno7", "weno9"})

_vp, _i, _d = _lib._vp, _lib._i, _lib._d
_lib.extra_signatures({
    f"fvx_advection_rk3_step{v}_{s}": [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _d, _d, _d,
                                      _vp]
    for s in ("f32", "f64") for v in ("", "_noalloc")
})


def method_codes(method: str) -> tuple[int, int]:
    if method in _LIMITERS:
        return 8, _LIMITERS[method]
    if method in _METHODS:
        return _METHODS[method], 0
    if method in _UNSUPPORTED:
        raise NotImplementedError(f"method {method!r} is outside this build's hot path")
    raise ValueError(f"Unknown method: {method!r}")


class Advection2D:
    """Advective tendency ``-div(h u)`` at T-points."""

    def __init__(self, grid: CartesianGrid2D, mask: Mask2D | None = None):
        self.grid = grid
        self.mask = mask

    def __call__(self, h, u, v, method: str = "upwind1"):
        m, lim = method_codes(method)
        h = _lib.as_field(h)
        u = _lib.as_field(u, h.dtype)
        v = _lib.as_field(v, h.dtype)
        if h.dim() != 2 or h.shape != u.shape or h.shape != v.shape:
            raise ValueError(f"h, u, v must share one 2-D shape; got {tuple(h.shape)}, "
                             f"{tuple(u.shape)}, {tuple(v.shape)}")
        ny, nx = h.shape
        tx = ty = mh = None
        if self.mask is not None:
            if tuple(self.mask.shape) != (ny, nx):
                raise ValueError(f"mask shape {self.mask.shape} != field shape {(ny, nx)}")
            mh = self.mask.device("h")
            if method in _MASK_DISPATCHABLE:
                tx, ty = self.mask.device("tier_x"), self.mask.device("tier_y")
        out = torch.empty_like(h)
        _lib.call(f"fvx_advection2d_{_lib.suffix(h)}", m, lim, h.data_ptr(), u.data_ptr(),
                  v.data_ptr(), out.data_ptr(), _lib.ptr(tx), _lib.ptr(ty), _lib.ptr(mh),
                  ny, nx, float(self.grid.dx), float(self.grid.dy), _lib.stream())
        return out

    def rk3_step(self, h, u, v, dt: float, method: str = "weno5", ng: int = 3):
        """One ``rk3_ssp_step`` of the doubly periodic tracer ``h``: before
        each tendency the ghost ring of width ``ng`` is filled periodically
        (``enforce_periodic``), and the result leaves with its ghosts filled;
        the same as ``enforce_periodic(rk3_ssp_step(h, lambda q:
        self(enforce_periodic(q, ng), u, v, method), dt), ng)`` as one native
        call (``fvx_advection_rk3_step_*``).  Returns a new array."""
        m, lim = method_codes(method)
        h = _lib.as_field(h).clone()
        u = _lib.as_field(u, h.dtype)
        v = _lib.as_field(v, h.dtype)
        if h.dim() != 2 or h.shape != u.shape or h.shape != v.shape:
            raise ValueError("h, u, v must share one 2-D shape")
        ny, nx = h.shape
        tx = ty = mh = None
        if self.mask is not None:
            mh = self.mask.device("h")
            if method in _MASK_DISPATCHABLE:
                tx, ty = self.mask.device("tier_x"), self.mask.device("tier_y")
        _lib.call(f"fvx_advection_rk3_step_{_lib.suffix(h)}", m, lim, h.data_ptr(), u.data_ptr(),
                  v.data_ptr(), _lib.ptr(tx), _lib.ptr(ty), _lib.ptr(mh), ny, nx, int(ng),
                  float(self.grid.dx), float(self.grid.dy), float(dt), _lib.stream())
        return h


_lib.extra_signatures({
    f"fvx_face_flux_{s}": [_i, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _vp] for s in ("f32", "f64")
})


class Reconstruction2D:
    """Face-value reconstruction times velocity: the east- (``*_x``) and
    north-face (``*_y``) fluxes of ``Reconstruction2D``
    (``advection/reconstruction.py:572-1307``), each one HIP kernel, zero
    ghost ring.  The ``*_masked`` methods pick the stencil per face from the
    upwind cell's tier of ``mask`` (upwind1 / 3-point / 5-point scheme,
    ``advection/flux.py:100-253``)."""

    def __init__(self, grid: CartesianGrid2D):
        self.grid = grid

    def _flux(self, method, dir_, h, vel, limiter="minmod", mask=None):
        m, lim = (8, _LIMITERS[limiter]) if method == "tvd" else (_RECON[method], 0)
        if method == "tvd" and limiter not in _LIMITERS:
            raise ValueError(f"Unknown limiter: {limiter!r}")
        h = _lib.as_field(h)
        vel = _lib.as_field(vel, h.dtype)
        if h.dim() != 2 or h.shape != vel.shape:
            raise ValueError("h and the face velocity must share one 2-D shape")
        ny, nx = h.shape
        tier = None
        if mask is not None:
            if tuple(mask.shape) != (ny, nx):
                raise ValueError(f"mask shape {mask.shape} != field shape {(ny, nx)}")
            tier = mask.device("tier_x" if dir_ == 0 else "tier_y")
        out = torch.empty_like(h)
        _lib.call(f"fvx_face_flux_{_lib.suffix(h)}", m, lim, dir_, h.data_ptr(), vel.data_ptr(),
                  out.data_ptr(), _lib.ptr(tier), ny, nx, _lib.stream())
        return out

    def tvd_x(self, h, u, limiter: str = "minmod"):
        return self._flux("tvd", 0, h, u, limiter)

    def tvd_y(self, h, v, limiter: str = "minmod"):
        return self._flux("tvd", 1, h, v, limiter)

    def tvd_x_masked(self, h, u, mask: Mask2D, limiter: str = "minmod"):
        return self._flux("tvd", 0, h, u, limiter, mask)

    def tvd_y_masked(self, h, v, mask: Mask2D, limiter: str = "minmod"):
        return self._flux("tvd", 1, h, v, limiter, mask)


# Reconstruction codes of the face-flux kernel (FVX_ADV_* in include/fvx.h).
_RECON = {"naive": 0, "upwind1": 1, "upwind2": 2, "upwind3": 3, "weno3": 4, "wenoz3": 5,
          "weno5": 6, "wenoz5": 7}


def _add_recon_methods():
    for name in _RECON:
        for d, axis in ((0, "x"), (1, "y")):
            def f(self, h, vel, _n=name, _d=d):
                return self._flux(_n, _d, h, vel)
            f.__name__ = f"{name}_{axis}"
            f.__doc__ = (f"{'East' if d == 0 else 'North'}-face {name} flux "
                         f"(reconstruction.py, Reconstruction2D.{name}_{axis}).")
            setattr(Reconstruction2D, f.__name__, f)
    for name in ("weno5", "wenoz5"):
        for d, axis in ((0, "x"), (1, "y")):
            def g(self, h, vel, mask, _n=name, _d=d):
                return self._flux(_n, _d, h, vel, mask=mask)
            g.__name__ = f"{name}_{axis}_masked"
            g.__doc__ = f"Tiered {name} flux (Reconstruction2D.{name}_{axis}_masked)."
            setattr(Reconstruction2D, g.__name__, g)
    for name in ("weno7", "weno9"):
        for axis in ("x", "y"):
            def n(self, *a, _n=name, **k):
                raise NotImplementedError(f"{_n} is outside this build's hot path")
            n.__name__ = f"{name}_{axis}"
            setattr(Reconstruction2D, n.__name__, n)


_add_recon_methods()
