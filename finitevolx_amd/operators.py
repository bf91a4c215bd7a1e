"""C-grid difference, divergence, interpolation and Jacobian operators.

Drop-in for the reference's ``Difference2D`` (``operators/difference.py:165-519``),
``Divergence2D`` / ``divergence_2d`` (``operators/divergence.py:17-163``), the
``Interpolation2D`` maps used by the QG diagnostics (``T_to_X``, ``V_to_T``,
``operators/interpolation.py:158-214``) and ``arakawa_jacobian``
(``operators/jacobian.py:22-131``).  Inputs are torch CUDA tensors (or NumPy
arrays, copied to the GPU); every call runs one HIP kernel.
"""

from __future__ import annotations

import torch

from finitevolx_amd import _lib
from finitevolx_amd.grid import CartesianGrid2D
from finitevolx_amd.mask import Mask2D

# Operator codes (FVX_OP_* in include/fvx.h)
OPS = {
    "diff_x_T_to_U": (0, "u"), "diff_y_T_to_V": (1, "v"),
    "diff_y_U_to_X": (2, "xy_corner_strict"), "diff_x_V_to_X": (3, "xy_corner_strict"),
    "diff_y_X_to_U": (4, "u"), "diff_x_X_to_V": (5, "v"),
    "diff_x_U_to_T": (6, "h"), "diff_y_V_to_T": (7, "h"),
    "divergence": (8, "h"), "noflux": (9, "h"), "curl": (10, "xy_corner_strict"),
    "laplacian": (11, "h"), "grad_perp_u": (12, "u"), "grad_perp_v": (13, "v"),
    "T_to_X": (14, "xy_corner_strict"), "V_to_T": (15, "h"),
}


def _stencil(name, a, b, dx, dy, mask: Mask2D | None):
    op, stagger = OPS[name]
    a = _lib.as_field(a)
    if a.dim() != 2:
        raise ValueError(f"expected a 2-D field, got shape {tuple(a.shape)}")
    b = None if b is None else _lib.as_field(b, a.dtype)
    if b is not None and b.shape != a.shape:
        raise ValueError(f"shape mismatch {tuple(a.shape)} vs {tuple(b.shape)}")
    m = None if mask is None else mask.device(stagger)
    if m is not None and tuple(m.shape) != tuple(a.shape):
        raise ValueError(f"mask shape {tuple(m.shape)} != field shape {tuple(a.shape)}")
    out = torch.empty_like(a)
    ny, nx = a.shape
    _lib.call(f"fvx_stencil2d_{_lib.suffix(a)}", op, a.data_ptr(), _lib.ptr(b), out.data_ptr(),
              ny, nx, float(dx), float(dy), _lib.ptr(m), _lib.stream())
    return out


class Difference2D:
    """Finite differences on a 2-D Arakawa C-grid (``operators/difference.py:165``).

    With ``mask`` set, outputs are multiplied by the stagger mask of the
    output point (T: ``h``, U: ``u``, V: ``v``, X: ``xy_corner_strict``).
    """

    def __init__(self, grid: CartesianGrid2D, mask: Mask2D | None = None):
        self.grid = grid
        self.mask = mask

    def _op(self, name, a, b=None):
        return _stencil(name, a, b, self.grid.dx, self.grid.dy, self.mask)

    def diff_x_T_to_U(self, h):
        return self._op("diff_x_T_to_U", h)

    def diff_y_T_to_V(self, h):
        return self._op("diff_y_T_to_V", h)

    def diff_y_U_to_X(self, u):
        return self._op("diff_y_U_to_X", u)

    def diff_x_V_to_X(self, v):
        return self._op("diff_x_V_to_X", v)

    def diff_y_X_to_U(self, q):
        return self._op("diff_y_X_to_U", q)

    def diff_x_X_to_V(self, q):
        return self._op("diff_x_X_to_V", q)

    def diff_x_U_to_T(self, u):
        return self._op("diff_x_U_to_T", u)

    def diff_y_V_to_T(self, v):
        return self._op("diff_y_V_to_T", v)

    def divergence(self, u, v):
        return self._op("divergence", u, v)

    def curl(self, u, v):
        return self._op("curl", u, v)

    def laplacian(self, h):
        return self._op("laplacian", h)

    def grad_perp(self, psi):
        """T-point psi -> (u, v) at U/V points (``difference.py:456-519``)."""
        return self._op("grad_perp_u", psi), self._op("grad_perp_v", psi)


def divergence_2d(u, v, dx: float, dy: float):
    """Functional divergence (``operators/divergence.py:17-60``)."""
    return _stencil("divergence", u, v, dx, dy, None)


class Divergence2D:
    """``operators/divergence.py:63-163``."""

    def __init__(self, grid: CartesianGrid2D, mask: Mask2D | None = None):
        self.grid = grid
        self.mask = mask
        self.diff = Difference2D(grid, mask)

    def __call__(self, u, v):
        return self.diff.divergence(u, v)

    def noflux(self, u, v):
        """Closed-basin variant: wall faces zeroed first (``divergence.py:138-163``)."""
        return _stencil("noflux", u, v, self.grid.dx, self.grid.dy, self.mask)


class Interpolation2D:
    """The T->X and V->T averages of ``operators/interpolation.py:104-214``."""

    def __init__(self, grid: CartesianGrid2D, mask: Mask2D | None = None):
        self.grid = grid
        self.mask = mask

    def T_to_X(self, h):
        return _stencil("T_to_X", h, None, self.grid.dx, self.grid.dy, self.mask)

    def V_to_T(self, v):
        return _stencil("V_to_T", v, None, self.grid.dx, self.grid.dy, self.mask)


def arakawa_jacobian(f, g, dx: float, dy: float):
    """Arakawa (1966) J(f, g) on the interior (``operators/jacobian.py:22-131``).

    Inputs ``(..., Ny, Nx)`` with a one-cell halo; output ``(..., Ny-2, Nx-2)``.
    """
    f = _lib.as_field(f)
    g = _lib.as_field(g, f.dtype)
    if f.shape != g.shape:
        raise ValueError(f"shape mismatch {tuple(f.shape)} vs {tuple(g.shape)}")
    if f.dim() < 2:
        raise ValueError("arakawa_jacobian needs at least 2-D inputs")
    ny, nx = f.shape[-2:]
    lead = f.shape[:-2]
    batch = 1
    for s in lead:
        batch *= s
    out = torch.empty((*lead, ny - 2, nx - 2), dtype=f.dtype, device=f.device)
    if batch == 0:
        return out
    _lib.call(f"fvx_arakawa_jacobian_{_lib.suffix(f)}", f.data_ptr(), g.data_ptr(),
              out.data_ptr(), batch, ny, nx, float(dx), float(dy), _lib.stream())
    return out


def interior(values, like, ghost: int = 1):
    """Embed ``values`` in zeros shaped like ``like`` (``operators/_ghost.py:9-28``)."""
    like = _lib.as_field(like)
    values = _lib.as_field(values, like.dtype)
    out = torch.zeros_like(like)
    out[..., ghost:-ghost, ghost:-ghost] = values
    return out
