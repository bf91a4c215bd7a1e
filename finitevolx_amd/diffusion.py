"""Harmonic and biharmonic diffusion (``diffusion/diffusion.py:70-574``)."""

from __future__ import annotations

import torch

from finitevolx_amd import _lib
from finitevolx_amd.grid import CartesianGrid2D
from finitevolx_amd.mask import Mask2D


def _kappa(kappa, like):
    if isinstance(kappa, (int, float)):
        return None, float(kappa)
    k = _lib.as_field(kappa, like.dtype)
    if k.dim() == 0:
        return None, float(k.item())
    if k.shape != like.shape:
        raise ValueError(f"kappa shape {tuple(k.shape)} != field shape {tuple(like.shape)}")
    return k, 0.0


def _masks(mask, names, shape):
    if mask is None:
        return [None] * len(names)
    if tuple(mask.shape) != tuple(shape):
        raise ValueError(f"mask shape {mask.shape} != field shape {tuple(shape)}")
    return [mask.device(n) for n in names]


def diffusion_2d(h, kappa, dx: float, dy: float):
    """Mask-free functional form (``diffusion.py:70-117``)."""
    return Diffusion2D(CartesianGrid2D(0, 0, 0.0, 0.0, dx, dy))(h, kappa)


class Diffusion2D:
    """Flux-form ``div(kappa grad h)`` with no-flux walls (``diffusion.py:209-329``)."""

    def __init__(self, grid: CartesianGrid2D, mask: Mask2D | None = None):
        self.grid = grid
        self.mask = mask

    def __call__(self, h, kappa):
        h = _lib.as_field(h)
        k, k0 = _kappa(kappa, h)
        mh, mu, mv = _masks(self.mask, ("h", "u", "v"), h.shape)
        out = torch.empty_like(h)
        ny, nx = h.shape
        _lib.call(f"fvx_diffusion2d_{_lib.suffix(h)}", h.data_ptr(), _lib.ptr(k), k0,
                  out.data_ptr(), ny, nx, float(self.grid.dx), float(self.grid.dy),
                  _lib.ptr(mh), _lib.ptr(mu), _lib.ptr(mv), _lib.stream())
        return out

    def fluxes(self, h, kappa):
        h = _lib.as_field(h)
        k, k0 = _kappa(kappa, h)
        mu, mv = _masks(self.mask, ("u", "v"), h.shape)
        fx, fy = torch.empty_like(h), torch.empty_like(h)
        ny, nx = h.shape
        _lib.call(f"fvx_diffusion2d_fluxes_{_lib.suffix(h)}", h.data_ptr(), _lib.ptr(k), k0,
                  fx.data_ptr(), fy.data_ptr(), ny, nx, float(self.grid.dx),
                  float(self.grid.dy), _lib.ptr(mu), _lib.ptr(mv), _lib.stream())
        return fx, fy


class BiharmonicDiffusion2D:
    """``-kappa * D(D(h))`` fused in one kernel (``diffusion.py:462-574``).

    The inner harmonic operator is unmasked; the mask multiplies the final
    tendency only, as in the reference.
    """

    def __init__(self, grid: CartesianGrid2D, mask: Mask2D | None = None):
        self.grid = grid
        self.mask = mask

    def __call__(self, h, kappa: float):
        h = _lib.as_field(h)
        (mh,) = _masks(self.mask, ("h",), h.shape)
        out = torch.empty_like(h)
        ny, nx = h.shape
        _lib.call(f"fvx_biharmonic2d_{_lib.suffix(h)}", h.data_ptr(), out.data_ptr(), ny, nx,
                  float(self.grid.dx), float(self.grid.dy), float(kappa), _lib.ptr(mh),
                  _lib.stream())
        return out


_lib.extra_signatures({
    f"fvx_momentum_advection2d_{s}": [_lib._vp, _lib._vp, _lib._vp, _lib._vp, _lib._i, _lib._vp,
                                      _lib._vp, _lib._vp, _lib._vp, _lib._i, _lib._i, _lib._d,
                                      _lib._d, _lib._vp]
    for s in ("f32", "f64")
})
_SCHEMES = {"energy": 0, "enstrophy": 1, "al": 2}


class MomentumAdvection2D:
    """Vector-invariant momentum advection (``diffusion/momentum.py:37-248``):
    ``(du, dv) = (+(zeta v)_u - dK/dx, -(zeta u)_v - dK/dy)`` with Sadourny's
    energy or enstrophy vorticity flux or the Arakawa-Lamb blend, as one HIP
    kernel (every intermediate stays in registers)."""

    def __init__(self, grid: CartesianGrid2D, mask: Mask2D | None = None):
        self.grid = grid
        self.mask = mask

    def __call__(self, u, v, scheme: str = "energy"):
        if scheme not in _SCHEMES:
            raise ValueError(f"Unknown scheme: {scheme!r}.  Choose 'energy', 'enstrophy', or 'al'.")
        u = _lib.as_field(u)
        v = _lib.as_field(v, u.dtype)
        if u.dim() != 2 or u.shape != v.shape:
            raise ValueError("u and v must share one 2-D shape")
        ny, nx = u.shape
        m = [None] * 4
        if self.mask is not None:
            m = [self.mask.device(k) for k in ("h", "u", "v", "xy_corner_strict")]
        du, dv = torch.empty_like(u), torch.empty_like(u)
        _lib.call(f"fvx_momentum_advection2d_{_lib.suffix(u)}", u.data_ptr(), v.data_ptr(),
                  du.data_ptr(), dv.data_ptr(), _SCHEMES[scheme], *[_lib.ptr(t) for t in m],
                  ny, nx, float(self.grid.dx), float(self.grid.dy), _lib.stream())
        return du, dv
