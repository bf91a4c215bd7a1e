"""Host-side land/ocean masks for the C-grid (setup, NumPy).

Mirrors the parts of ``Mask2D`` (reference ``mask/cartesian.py:345-739``)
that the hot path consumes: the staggered masks ``h, u, v, xy_corner,
xy_corner_strict`` (trailing-pad pooling, ``mask/utils.py:42-140``), the
directional stencil capability (``mask/base.py:60-105`` with
``count_contiguous``, ``mask/utils.py:385-468``) and the mutually exclusive
adaptive stencil tiers (``get_adaptive_masks``, ``mask/cartesian.py:495-553``).

For the GPU the tiers are stored as one uint8 per cell and direction holding
the largest usable stencil size in {0, 2, 4, 6}; device copies are cached.
"""

from __future__ import annotations

import numpy as np

from finitevolx_amd import _lib


def _pool_trailing(h: np.ndarray, ky: int, kx: int, threshold: float) -> np.ndarray:
    """Mean over the trailing ky x kx window (zero past the end) > threshold."""
    ny, nx = h.shape
    p = np.zeros((ny + ky - 1, nx + kx - 1), dtype=np.float32)
    p[:ny, :nx] = h
    acc = np.zeros((ny, nx), dtype=np.float32)
    for oy in range(ky):
        for ox in range(kx):
            acc += p[oy:oy + ny, ox:ox + nx]
    return acc / (ky * kx) > threshold


def _runs(h: np.ndarray, axis: int, forward: bool) -> np.ndarray:
    """Count of consecutive wet cells starting at each cell (inclusive)."""
    a = np.moveaxis(h.astype(np.int32), axis, 0)
    c = np.zeros_like(a)
    n = a.shape[0]
    rng = range(n - 1, -1, -1) if forward else range(n)
    prev = None
    for k in rng:
        c[k] = a[k] if prev is None else a[k] * (1 + c[prev])
        prev = k
    return np.moveaxis(c, 0, axis)


class StencilCapability2D:
    """Directional contiguous-wet counts (``mask/base.py:60-105``)."""

    def __init__(self, h: np.ndarray):
        self.x_pos = _runs(h, 1, True)
        self.x_neg = _runs(h, 1, False)
        self.y_pos = _runs(h, 0, True)
        self.y_neg = _runs(h, 0, False)


class Mask2D:
    """Staggered wet masks plus adaptive-stencil tiers for one 2-D domain."""

    def __init__(self, h: np.ndarray):
        h = np.asarray(h, dtype=bool)
        if h.ndim != 2:
            raise ValueError(f"mask must be 2-D, got shape {h.shape}")
        self.h = h
        self.u = _pool_trailing(h, 1, 2, 3.0 / 4.0)
        self.v = _pool_trailing(h, 2, 1, 3.0 / 4.0)
        self.xy_corner = _pool_trailing(h, 2, 2, 1.0 / 8.0)
        self.xy_corner_strict = _pool_trailing(h, 2, 2, 7.0 / 8.0)
        self.stencil_capability = StencilCapability2D(h)
        self._dev: dict[str, object] = {}

    # --- factories ---------------------------------------------------------
    @classmethod
    def from_mask(cls, mask_hgrid) -> "Mask2D":
        """``Mask2D.from_mask`` (``mask/cartesian.py:584-739``)."""
        return cls(np.asarray(mask_hgrid) != 0)

    @classmethod
    def from_dimensions(cls, ny: int, nx: int) -> "Mask2D":
        """All-ocean mask of total shape ``(ny, nx)``."""
        return cls(np.ones((ny, nx), dtype=bool))

    @property
    def shape(self):
        return self.h.shape

    # --- adaptive stencil tiers --------------------------------------------
    def max_stencil(self, direction: str, stencil_sizes=(2, 4, 6)) -> np.ndarray:
        """Largest usable stencil per cell (0 where none), as uint8."""
        sc = self.stencil_capability
        if direction == "x":
            cp, cn = sc.x_pos, sc.x_neg
        elif direction == "y":
            cp, cn = sc.y_pos, sc.y_neg
        else:
            raise ValueError(f"direction must be 'x' or 'y', got {direction!r}")
        out = np.zeros(self.h.shape, dtype=np.uint8)
        for s in sorted(stencil_sizes):
            out = np.where((cp >= s // 2) & (cn >= s // 2), np.uint8(s), out)
        return out

    def get_adaptive_masks(self, direction: str = "x", stencil_sizes=(2, 4, 6, 8, 10)):
        """Mutually exclusive boolean tiers (``mask/cartesian.py:495-553``)."""
        m = self.max_stencil(direction, stencil_sizes)
        return {s: (m == s) for s in stencil_sizes}

    # --- device copies -----------------------------------------------------
    def device(self, name: str):
        """uint8 CUDA tensor of a stagger mask or tier map (cached)."""
        t = self._dev.get(name)
        if t is None:
            import torch

            if name in ("tier_x", "tier_y"):
                arr = self.max_stencil(name[-1], (2, 4, 6))
            else:
                arr = getattr(self, name).astype(np.uint8)
            t = torch.from_numpy(np.ascontiguousarray(arr)).to(_lib.require_gpu())
            self._dev[name] = t
        return t
