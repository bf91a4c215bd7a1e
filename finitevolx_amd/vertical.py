"""Vertical-mode decomposition for multi-layer QG (``vertical/vertical_modes.py``).

The coupling matrix and its eigendecomposition are tiny host computations
(as in the reference); the per-cell layer <-> mode transforms run on the GPU
in one pass (``fvx_layer_mix``).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from finitevolx_amd import _lib

_lib.extra_signatures({
    f"fvx_layer_mix_{s}": [_lib._vp, _lib._vp, _lib._vp, _lib._i, C.c_int64, _lib._vp]
    for s in ("f32", "f64")
})


def build_coupling_matrix(H, g_prime) -> np.ndarray:
    """Tridiagonal stratification matrix A (``vertical_modes.py:50-121``)."""
    H = np.asarray(H, dtype=np.float64)
    g = np.asarray(g_prime, dtype=np.float64)
    if H.shape != g.shape or H.ndim != 1:
        raise ValueError("H and g_prime must be 1-D arrays of equal length")
    nl = H.size
    A = np.zeros((nl, nl))
    if nl == 1:
        A[0, 0] = 1.0 / (H[0] * g[0])
        return A
    A[0, 0] = 1.0 / (H[0] * g[0]) + 1.0 / (H[0] * g[1])
    A[0, 1] = -1.0 / (H[0] * g[1])
    for i in range(1, nl - 1):
        A[i, i - 1] = -1.0 / (H[i] * g[i])
        A[i, i] = (1.0 / H[i]) * (1.0 / g[i + 1] + 1.0 / g[i])
        A[i, i + 1] = -1.0 / (H[i] * g[i + 1])
    A[-1, -1] = 1.0 / (H[-1] * g[-1])
    A[-1, -2] = -1.0 / (H[-1] * g[-1])
    return A


def decompose_vertical_modes(A, f0: float):
    """Rossby radii and transforms (``vertical_modes.py:124-190``).

    ``jnp.linalg.eigh`` symmetrises its input, so ``(A + A^T)/2`` is
    decomposed here; ``Cl2m = R^T``, ``Cm2l = R``; non-positive eigenvalues
    give an infinite radius.
    """
    A = np.asarray(A, dtype=np.float64)
    lam, R = np.linalg.eigh(0.5 * (A + A.T))
    pos = lam > 0
    radii = np.where(pos, 1.0 / (abs(f0) * np.sqrt(np.where(pos, lam, 1.0))), np.inf)
    return radii, R.T.copy(), R.copy()


def _mix(field, Cmat):
    f = _lib.as_field(field)
    nl = f.shape[0]
    Cm = np.ascontiguousarray(np.asarray(Cmat, dtype=np.float64))
    if Cm.shape != (nl, nl):
        raise ValueError(f"transform shape {Cm.shape} does not match {nl} layers")
    flat = f.reshape(nl, -1).contiguous()
    out = torch.empty_like(flat)
    _lib.call(f"fvx_layer_mix_{_lib.suffix(f)}", flat.data_ptr(), out.data_ptr(), Cm.ctypes.data,
              nl, flat.shape[1], _lib.stream())
    return out.reshape(f.shape)


def layer_to_mode(field, Cl2m):
    """``mode = Cl2m @ layer`` along the leading axis (``vertical_modes.py:193-227``)."""
    return _mix(field, Cl2m)


def mode_to_layer(field, Cm2l):
    """``layer = Cm2l @ mode`` along the leading axis (``vertical_modes.py:230-263``)."""
    return _mix(field, Cm2l)
