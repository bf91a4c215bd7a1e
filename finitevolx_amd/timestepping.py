"""Explicit Runge-Kutta stepping (``timestepping/explicit_rk.py``):
``euler_step`` (:22-40), ``heun_step`` (:43-71), ``rk3_ssp_step`` (:74-121)
and ``rk4_step`` (:124-168).

The stage combinations run as one fused HIP kernel per leaf; ``state`` may
be a tensor or a (nested) tuple / list / dict of tensors, like the
reference's pytrees.
"""

from __future__ import annotations

import torch

from finitevolx_amd import _lib

_lib.extra_signatures({
    **{f"fvx_lincomb3_{s}": [_lib.C.c_int64, _lib._d, _lib._vp, _lib._d, _lib._vp, _lib._d,
                             _lib._vp, _lib._vp, _lib._vp] for s in ("f32", "f64")},
    **{f"fvx_lincomb4_{s}": [_lib.C.c_int64, _lib._d, _lib._vp, _lib._d, _lib._vp, _lib._d,
                             _lib._vp, _lib._d, _lib._vp, _lib._vp, _lib._vp]
       for s in ("f32", "f64")},
})


def _tree_map(fn, *trees):
    t0 = trees[0]
    if isinstance(t0, dict):
        return {k: _tree_map(fn, *[t[k] for t in trees]) for k in t0}
    if isinstance(t0, (tuple, list)):
        return type(t0)(_tree_map(fn, *xs) for xs in zip(*trees))
    return fn(*trees)


def lincomb(a, x, b=0.0, y=None, c=0.0, z=None, out=None, d=0.0, w=None):
    """``a*x + b*y + c*z + d*w`` on device tensors in one kernel (into ``out``
    when given: a contiguous tensor of x's shape and type, which may be one of
    the inputs -- each element is read before it is written)."""
    x = _lib.as_field(x)
    y = None if y is None else _lib.as_field(y, x.dtype)
    z = None if z is None else _lib.as_field(z, x.dtype)
    w = None if w is None else _lib.as_field(w, x.dtype)
    for t in (y, z, w):
        if t is not None and t.shape != x.shape:
            raise ValueError("shape mismatch in lincomb")
    if out is None:
        out = torch.empty_like(x)
    elif out.shape != x.shape or out.dtype != x.dtype or not out.is_contiguous():
        raise ValueError("out must be a contiguous tensor like x")
    if w is None:
        _lib.call(f"fvx_lincomb3_{_lib.suffix(x)}", x.numel(), float(a), x.data_ptr(), float(b),
                  _lib.ptr(y), float(c), _lib.ptr(z), out.data_ptr(), _lib.stream())
    else:
        _lib.call(f"fvx_lincomb4_{_lib.suffix(x)}", x.numel(), float(a), x.data_ptr(), float(b),
                  _lib.ptr(y), float(c), _lib.ptr(z), float(d), w.data_ptr(), out.data_ptr(),
                  _lib.stream())
    return out


def rk3_ssp_step(state, rhs_fn, dt: float):
    """Shu-Osher SSP-RK3: three RHS calls, fused stage combinations."""
    k1 = rhs_fn(state)
    y1 = _tree_map(lambda y, f: lincomb(1.0, y, dt, f), state, k1)
    k2 = rhs_fn(y1)
    y2 = _tree_map(lambda y, a, f: lincomb(0.75, y, 0.25, a, 0.25 * dt, f), state, y1, k2)
    k3 = rhs_fn(y2)
    return _tree_map(lambda y, a, f: lincomb(1.0 / 3.0, y, 2.0 / 3.0, a, 2.0 / 3.0 * dt, f),
                     state, y2, k3)


def euler_step(state, rhs_fn, dt: float):
    """Forward Euler: ``y + dt f(y)``."""
    k1 = rhs_fn(state)
    return _tree_map(lambda y, f: lincomb(1.0, y, dt, f), state, k1)


def heun_step(state, rhs_fn, dt: float):
    """Heun (RK2): ``y* = y + dt k1``, ``y + dt/2 (k1 + k2)`` with ``k2 = f(y*)``."""
    k1 = rhs_fn(state)
    ys = _tree_map(lambda y, f: lincomb(1.0, y, dt, f), state, k1)
    k2 = rhs_fn(ys)
    return _tree_map(lambda y, f1, f2: lincomb(1.0, y, 0.5 * dt, f1, 0.5 * dt, f2), state, k1, k2)


def rk4_step(state, rhs_fn, dt: float):
    """Classical RK4: ``y + dt/6 (k1 + 2 k2 + 2 k3 + k4)``."""
    k1 = rhs_fn(state)
    k2 = rhs_fn(_tree_map(lambda y, f: lincomb(1.0, y, 0.5 * dt, f), state, k1))
    k3 = rhs_fn(_tree_map(lambda y, f: lincomb(1.0, y, 0.5 * dt, f), state, k2))
    k4 = rhs_fn(_tree_map(lambda y, f: lincomb(1.0, y, dt, f), state, k3))
    part = _tree_map(lambda y, a, b: lincomb(1.0, y, dt / 6.0, a, dt / 3.0, b), state, k1, k2)
    return _tree_map(lambda y, a, b: lincomb(1.0, y, dt / 3.0, a, dt / 6.0, b), part, k3, k4)
