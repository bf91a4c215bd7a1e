"""NumPy float64 restatement of the finitevolX QG hot path (the oracle).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, never by ``finitevolx_amd``.

Each function restates the reference algorithm and cites the reference
file:line it follows (paths relative to the finitevolX 0.0.40 tree).
Array convention (reference ``grid/cartesian.py:56-80``): row-major
``[Ny, Nx] = [j, i]`` including a one-cell ghost ring; T[j,i] centre,
U[j,i] east face, V[j,i] north face, X[j,i] NE corner.

How each function is pinned (each check would fail on a plausible planted
mistake; ``tests/test_oracle_pins.py`` plants one per family):

* stencils, Difference2D, Divergence2D, Interpolation T_to_X / V_to_T,
  Reconstruction/Advection2D (upwind1, weno5; masked and unmasked),
  Diffusion2D, BiharmonicDiffusion2D: the reference's own golden ``.npz``
  files (copied to ``tests/golden/``) at rtol 1e-12 / atol 1e-14
  (``tests/test_oracle_golden.py``).
* WENO3/WENO5/WENO-Z point kernels: constant and linear exactness, the
  left/right mirror identity, boundedness at a step, upwind selection, and
  the formal order of accuracy on cell averages (``test_oracle_pins.py``,
  after reference ``tests/test_weno.py``).
* TVD limiters: the reference's closed-form values
  (``tests/test_limiters.py:11-106``), phi(1/r) = phi(r)/r and the TVD
  region.
* naive/upwind1/upwind2/upwind3/weno3/wenoz5/TVD face fluxes: linear
  exactness, order of accuracy, mirror symmetry under x -> -x, the wall
  fallbacks; Advection2D: constant fields give zero tendency for both flow
  signs, masked all-ocean equals unmasked, periodic rk3_ssp_step
  convergence order (reference ``tests/test_advection.py:116-628``).
* arakawa_jacobian: J(x,y)=1, J(xy,x^2)=-2x^2, antisymmetry, J(f,f)=0,
  constant arguments, periodic conservation of the integral, energy and
  enstrophy (reference ``tests/test_jacobian.py:93-230``).
* masked_laplacian and the multigrid operator: assembled as sparse
  matrices and checked for symmetry, against hand stencils, and against
  direct sparse solves; transfers exact on constants
  (reference ``tests/test_multigrid.py:117-301``; ``test_oracle_solvers.py``).
* PCG / flexible PCG: the converged solution equals a direct sparse solve.
  The reference delegates CG to lineax 0.1.0 (not vendored); its iterate
  sequence and stopping rule are parity unpinned, see ``pcg``.
* rk3_ssp_step: its stability polynomial 1 + z + z^2/2 + z^3/6 on
  y' = lambda y, third-order convergence on y' = -y^2, and the reference's
  1e-6 accuracy at dt = 0.01 (``tests/test_timestepping.py:60-106``).
* enforce_periodic: equal to numpy's wrap padding of the interior.
* vertical modes: closed forms of reference ``tests/test_vertical_modes.py``.
* QG step (build-defined composition, no single reference function):
  pinned through its parts and against a direct-solve composition.
* multi-layer QG step (composition of the vertical-mode API): the inverted
  streamfunction satisfies the textbook layer equations
  lap psi - f0^2 A psi = q when every eigenvalue is positive (equal layer
  thicknesses); one layer reduces to the 1.5-layer step with
  Ld^2 = H g / f0^2; the tracer velocity is exactly uniform for a linear
  streamfunction and discretely non-divergent; forcing enters the top layer
  only (``test_oracle_pins.py``).
"""

from __future__ import annotations

import numpy as np

# ---------------------------------------------------------------------------
# Ghost ring and raw stencils
# ---------------------------------------------------------------------------


def interior(values, like_shape, ghost=1):
    """Embed ``values`` in zeros of ``like_shape``.

    Follows ``operators/_ghost.py:9-28``.
    """
    out = np.zeros(like_shape, dtype=np.float64)
    out[ghost:-ghost, ghost:-ghost] = values
    return out


def diff_x_fwd(h):
    """``operators/stencils.py:84-107``: h[j,i+1] - h[j,i] on the interior."""
    return h[1:-1, 2:] - h[1:-1, 1:-1]


def diff_y_fwd(h):
    """``operators/stencils.py:110-132``."""
    return h[2:, 1:-1] - h[1:-1, 1:-1]


def diff_x_bwd(h):
    """``operators/stencils.py:135-157``."""
    return h[1:-1, 1:-1] - h[1:-1, :-2]


def diff_y_bwd(h):
    """``operators/stencils.py:160-181``."""
    return h[1:-1, 1:-1] - h[:-2, 1:-1]


# ---------------------------------------------------------------------------
# Masks (host-side setup)
# ---------------------------------------------------------------------------


def pool_bool_trailing(arr, kernel, threshold):
    """Windowed mean with trailing zero pad, compared with ``> threshold``.

    Follows ``mask/utils.py:42-140`` (``direction='trailing'``).
    """
    a = np.asarray(arr, dtype=np.float64)
    ky, kx = kernel
    ny, nx = a.shape
    pad = np.zeros((ny + ky - 1, nx + kx - 1))
    pad[:ny, :nx] = a
    acc = np.zeros((ny, nx))
    for oy in range(ky):
        for ox in range(kx):
            acc += pad[oy:oy + ny, ox:ox + nx]
    return (acc / (ky * kx)) > threshold


def count_contiguous(arr, axis, forward):
    """Contiguous wet-cell counts along an axis (``mask/utils.py:385-468``)."""
    a = np.moveaxis(np.asarray(arr, dtype=np.int32), axis, 0)
    c = np.zeros_like(a)
    n = a.shape[0]
    if forward:
        c[-1] = a[-1]
        for i in range(n - 2, -1, -1):
            c[i] = a[i] * (1 + c[i + 1])
    else:
        c[0] = a[0]
        for i in range(1, n):
            c[i] = a[i] * (1 + c[i - 1])
    return np.moveaxis(c, 0, axis)


class Mask2D:
    """The staggered masks the hot path consumes.

    Follows ``Mask2D.from_mask`` (``mask/cartesian.py:584-739``) for
    ``h, u, v, xy_corner, xy_corner_strict`` and the stencil capability
    (``mask/base.py:60-105``).
    """

    def __init__(self, h):
        h = np.asarray(h, dtype=bool)
        self.h = h
        self.u = pool_bool_trailing(h, (1, 2), 3.0 / 4.0)
        self.v = pool_bool_trailing(h, (2, 1), 3.0 / 4.0)
        self.xy_corner = pool_bool_trailing(h, (2, 2), 1.0 / 8.0)
        self.xy_corner_strict = pool_bool_trailing(h, (2, 2), 7.0 / 8.0)
        self.x_pos = count_contiguous(h, 1, True)
        self.x_neg = count_contiguous(h, 1, False)
        self.y_pos = count_contiguous(h, 0, True)
        self.y_neg = count_contiguous(h, 0, False)

    def get_adaptive_masks(self, direction="x", stencil_sizes=(2, 4, 6)):
        """``mask/cartesian.py:495-553``: mutually exclusive stencil tiers."""
        if direction == "x":
            cp, cn = self.x_pos, self.x_neg
        elif direction == "y":
            cp, cn = self.y_pos, self.y_neg
        else:
            raise ValueError(f"direction must be 'x' or 'y', got {direction!r}")
        max_s = np.zeros(self.h.shape, dtype=np.int32)
        for s in sorted(stencil_sizes):
            hw = s // 2
            max_s = np.where((cp >= hw) & (cn >= hw), s, max_s)
        return {s: (max_s == s) for s in stencil_sizes}


def narrow_mask_hierarchy(hier, sizes):
    """``advection/flux.py:41-97``: OR-fold tiers >= the largest target."""
    sizes = sorted(sizes)
    out = {s: hier[s] for s in sizes[:-1]}
    big = sizes[-1]
    fold = None
    for s, m in hier.items():
        if s >= big:
            fold = m if fold is None else (fold | m)
    out[big] = fold
    return out


# ---------------------------------------------------------------------------
# Difference2D / Divergence2D / Interpolation2D
# ---------------------------------------------------------------------------


def _apply_mask(out, mask_arr):
    return out if mask_arr is None else out * mask_arr


def diff_x_T_to_U(h, dx, mask=None):
    """``operators/difference.py:194-213``."""
    return _apply_mask(interior(diff_x_fwd(h) / dx, h.shape), None if mask is None else mask.u)


def diff_y_T_to_V(h, dy, mask=None):
    """``operators/difference.py:215-234``."""
    return _apply_mask(interior(diff_y_fwd(h) / dy, h.shape), None if mask is None else mask.v)


def diff_y_U_to_X(u, dy, mask=None):
    """``operators/difference.py:236-256``."""
    return _apply_mask(interior(diff_y_fwd(u) / dy, u.shape),
                       None if mask is None else mask.xy_corner_strict)


def diff_x_V_to_X(v, dx, mask=None):
    """``operators/difference.py:258-278``."""
    return _apply_mask(interior(diff_x_fwd(v) / dx, v.shape),
                       None if mask is None else mask.xy_corner_strict)


def diff_y_X_to_U(q, dy, mask=None):
    """``operators/difference.py:280-299``."""
    return _apply_mask(interior(diff_y_bwd(q) / dy, q.shape), None if mask is None else mask.u)


def diff_x_X_to_V(q, dx, mask=None):
    """``operators/difference.py:301-320``."""
    return _apply_mask(interior(diff_x_bwd(q) / dx, q.shape), None if mask is None else mask.v)


def diff_x_U_to_T(u, dx, mask=None):
    """``operators/difference.py:326-345``."""
    return _apply_mask(interior(diff_x_bwd(u) / dx, u.shape), None if mask is None else mask.h)


def diff_y_V_to_T(v, dy, mask=None):
    """``operators/difference.py:347-366``."""
    return _apply_mask(interior(diff_y_bwd(v) / dy, v.shape), None if mask is None else mask.h)


def divergence(u, v, dx, dy, mask=None):
    """``operators/difference.py:62-74`` and ``:372-399``."""
    out = interior(diff_x_bwd(u) / dx + diff_y_bwd(v) / dy, u.shape)
    return _apply_mask(out, None if mask is None else mask.h)


def divergence_noflux(u, v, dx, dy, mask=None):
    """``operators/divergence.py:138-163``: zero the four wall faces first."""
    ub = u.copy()
    vb = v.copy()
    ub[:, 0] = 0.0
    ub[:, -2] = 0.0
    vb[0, :] = 0.0
    vb[-2, :] = 0.0
    return divergence(ub, vb, dx, dy, mask)


def curl(u, v, dx, dy, mask=None):
    """``operators/difference.py:47-59`` and ``:401-429``."""
    out = interior(diff_x_fwd(v) / dx - diff_y_fwd(u) / dy, u.shape)
    return _apply_mask(out, None if mask is None else mask.xy_corner_strict)


def laplacian(h, dx, dy, mask=None):
    """``operators/difference.py:431-454``: 5-point Laplacian on [1:-1,1:-1]."""
    d2x = (diff_x_fwd(h) - diff_x_bwd(h)) / dx**2
    d2y = (diff_y_fwd(h) - diff_y_bwd(h)) / dy**2
    return _apply_mask(interior(d2x + d2y, h.shape), None if mask is None else mask.h)


def grad_perp(psi, dx, dy, mask=None):
    """``operators/difference.py:456-519``: T-point psi -> (u, v)."""
    u = interior(-(psi[2:, 1:-1] + psi[2:, 2:] - psi[:-2, 1:-1] - psi[:-2, 2:]) / (4.0 * dy),
                 psi.shape)
    v = interior((psi[1:-1, 2:] + psi[2:, 2:] - psi[1:-1, :-2] - psi[2:, :-2]) / (4.0 * dx),
                 psi.shape)
    if mask is not None:
        u = u * mask.u
        v = v * mask.v
    return u, v


def T_to_X(h, mask=None):
    """``operators/interpolation.py:158-168`` with ``stencils.py:449``."""
    out = interior(0.25 * (h[1:-1, 1:-1] + h[1:-1, 2:] + h[2:, 1:-1] + h[2:, 2:]), h.shape)
    return _apply_mask(out, None if mask is None else mask.xy_corner_strict)


def V_to_T(v, mask=None):
    """``operators/interpolation.py:206-214`` with ``stencils.py:424``."""
    out = interior(0.5 * (v[1:-1, 1:-1] + v[:-2, 1:-1]), v.shape)
    return _apply_mask(out, None if mask is None else mask.h)


def _avg(h, a, b, c=None, d=None):
    out = 0.5 * (h[a] + h[b]) if c is None else 0.25 * (h[a] + h[b] + h[c] + h[d])
    return out


_C = (slice(1, -1), slice(1, -1))
_E = (slice(1, -1), slice(2, None))
_W = (slice(1, -1), slice(None, -2))
_N = (slice(2, None), slice(1, -1))
_S = (slice(None, -2), slice(1, -1))
_NW = (slice(2, None), slice(None, -2))
_SE = (slice(None, -2), slice(2, None))


def _interp(out, h, stagger, mask):
    o = interior(out, h.shape)
    return o if mask is None else o * getattr(mask, stagger)


def U_to_T(u, mask=None):
    """``operators/interpolation.py:196-204`` (avg_x_bwd)."""
    return _interp(_avg(u, _C, _W), u, "h", mask)


def X_to_U(q, mask=None):
    """``operators/interpolation.py:172-180`` (avg_y_bwd)."""
    return _interp(_avg(q, _C, _S), q, "u", mask)


def X_to_V(q, mask=None):
    """``operators/interpolation.py:182-190`` (avg_x_bwd)."""
    return _interp(_avg(q, _C, _W), q, "v", mask)


def U_to_X(u, mask=None):
    """``operators/interpolation.py:231-239`` (avg_y_fwd)."""
    return _interp(_avg(u, _C, _N), u, "xy_corner_strict", mask)


def V_to_X(v, mask=None):
    """``operators/interpolation.py:241-249`` (avg_x_fwd)."""
    return _interp(_avg(v, _C, _E), v, "xy_corner_strict", mask)


def U_to_V(u, mask=None):
    """``operators/interpolation.py:255-264`` (avg_xbwd_yfwd)."""
    return _interp(_avg(u, _C, _N, _W, _NW), u, "v", mask)


def V_to_U(v, mask=None):
    """``operators/interpolation.py:266-275`` (avg_xfwd_ybwd)."""
    return _interp(_avg(v, _C, _S, _E, _SE), v, "u", mask)


def momentum_advection_2d(u, v, dx, dy, scheme="energy", mask=None):
    """``MomentumAdvection2D.__call__`` (``diffusion/momentum.py:97-248``):
    (du, dv) = (+(zeta v)_u - dK/dx, -(zeta u)_v - dK/dy) on [2:-2, 2:-2]."""
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    zeta = curl(u, v, dx, dy, mask)
    uT, vT = U_to_T(u, mask), V_to_T(v, mask)
    K = interior(0.5 * (uT[1:-1, 1:-1] ** 2 + vT[1:-1, 1:-1] ** 2), u.shape)
    dKdx, dKdy = diff_x_T_to_U(K, dx, mask), diff_y_T_to_V(K, dy, mask)

    def energy():
        return (interior((X_to_U(zeta, mask) * V_to_U(v, mask))[1:-1, 1:-1], u.shape),
                interior((X_to_V(zeta, mask) * U_to_V(u, mask))[1:-1, 1:-1], u.shape))

    def enstrophy():
        zvq = interior((zeta * V_to_X(v, mask))[1:-1, 1:-1], u.shape)
        zuq = interior((zeta * U_to_X(u, mask))[1:-1, 1:-1], u.shape)
        return X_to_U(zvq, mask), X_to_V(zuq, mask)

    if scheme == "energy":
        zv_u, zu_v = energy()
    elif scheme == "enstrophy":
        zv_u, zu_v = enstrophy()
    elif scheme == "al":
        (ea, eb), (sa, sb) = energy(), enstrophy()
        zv_u = interior(ea[1:-1, 1:-1] / 3.0 + (2.0 / 3.0) * sa[1:-1, 1:-1], u.shape)
        zu_v = interior(eb[1:-1, 1:-1] / 3.0 + (2.0 / 3.0) * sb[1:-1, 1:-1], u.shape)
    else:
        raise ValueError(f"Unknown scheme: {scheme!r}.  Choose 'energy', 'enstrophy', or 'al'.")
    du = interior(zv_u[2:-2, 2:-2] - dKdx[2:-2, 2:-2], u.shape, ghost=2)
    dv = interior(-zu_v[2:-2, 2:-2] - dKdy[2:-2, 2:-2], u.shape, ghost=2)
    if mask is not None:
        du, dv = du * mask.u, dv * mask.v
    return du, dv


# ---------------------------------------------------------------------------
# WENO point kernels and TVD limiters
# ---------------------------------------------------------------------------


def weno_3pts(qm, q0, qp):
    """``advection/weno.py:297-320`` (Jiang & Shu 1996, eps 1e-8)."""
    eps = 1e-8
    qi1 = -0.5 * qm + 1.5 * q0
    qi2 = 0.5 * (q0 + qp)
    b1 = (q0 - qm) ** 2
    b2 = (qp - q0) ** 2
    w1 = (1.0 / 3.0) / (b1 + eps) ** 2
    w2 = (2.0 / 3.0) / (b2 + eps) ** 2
    return (w1 * qi1 + w2 * qi2) / (w1 + w2)


def weno_3pts_right(q0, qp, qpp):
    """``advection/weno.py:413-435``."""
    eps = 1e-8
    qi1 = 0.5 * (q0 + qp)
    qi2 = 1.5 * qp - 0.5 * qpp
    b1 = (qp - q0) ** 2
    b2 = (qpp - qp) ** 2
    w1 = (2.0 / 3.0) / (b1 + eps) ** 2
    w2 = (1.0 / 3.0) / (b2 + eps) ** 2
    return (w1 * qi1 + w2 * qi2) / (w1 + w2)


def wenoz_3pts(qm, q0, qp):
    """``advection/weno.py:323-347`` (Borges et al. 2008, eps 1e-14)."""
    eps = 1e-14
    qi1 = -0.5 * qm + 1.5 * q0
    qi2 = 0.5 * (q0 + qp)
    b1 = (q0 - qm) ** 2
    b2 = (qp - q0) ** 2
    tau = np.abs(b2 - b1)
    w1 = (1.0 / 3.0) * (1.0 + tau / (b1 + eps))
    w2 = (2.0 / 3.0) * (1.0 + tau / (b2 + eps))
    return (w1 * qi1 + w2 * qi2) / (w1 + w2)


def wenoz_3pts_right(q0, qp, qpp):
    """``advection/weno.py:438-459``."""
    eps = 1e-14
    qi1 = 0.5 * (q0 + qp)
    qi2 = 1.5 * qp - 0.5 * qpp
    b1 = (qp - q0) ** 2
    b2 = (qpp - qp) ** 2
    tau = np.abs(b2 - b1)
    w1 = (2.0 / 3.0) * (1.0 + tau / (b1 + eps))
    w2 = (1.0 / 3.0) * (1.0 + tau / (b2 + eps))
    return (w1 * qi1 + w2 * qi2) / (w1 + w2)


def _weno5_parts(qmm, qm, q0, qp, qpp):
    qi1 = 1.0 / 3.0 * qmm - 7.0 / 6.0 * qm + 11.0 / 6.0 * q0
    qi2 = -1.0 / 6.0 * qm + 5.0 / 6.0 * q0 + 1.0 / 3.0 * qp
    qi3 = 1.0 / 3.0 * q0 + 5.0 / 6.0 * qp - 1.0 / 6.0 * qpp
    k1, k2 = 13.0 / 12.0, 0.25
    b1 = k1 * (qmm - 2 * qm + q0) ** 2 + k2 * (qmm - 4 * qm + 3 * q0) ** 2
    b2 = k1 * (qm - 2 * q0 + qp) ** 2 + k2 * (qm - qp) ** 2
    b3 = k1 * (q0 - 2 * qp + qpp) ** 2 + k2 * (3 * q0 - 4 * qp + qpp) ** 2
    return qi1, qi2, qi3, b1, b2, b3


def weno_5pts(qmm, qm, q0, qp, qpp):
    """``advection/weno.py:350-376`` (g = 0.1, 0.6, 0.3; eps 1e-8)."""
    eps = 1e-8
    qi1, qi2, qi3, b1, b2, b3 = _weno5_parts(qmm, qm, q0, qp, qpp)
    w1 = 0.1 / (b1 + eps) ** 2
    w2 = 0.6 / (b2 + eps) ** 2
    w3 = 0.3 / (b3 + eps) ** 2
    return (w1 * qi1 + w2 * qi2 + w3 * qi3) / (w1 + w2 + w3)


def _weno5_right_parts(qm, q0, qp, qpp, qppp):
    qi1 = 1.0 / 3.0 * qppp - 7.0 / 6.0 * qpp + 11.0 / 6.0 * qp
    qi2 = -1.0 / 6.0 * qpp + 5.0 / 6.0 * qp + 1.0 / 3.0 * q0
    qi3 = 1.0 / 3.0 * qp + 5.0 / 6.0 * q0 - 1.0 / 6.0 * qm
    k1, k2 = 13.0 / 12.0, 0.25
    b1 = k1 * (qp - 2 * qpp + qppp) ** 2 + k2 * (3 * qp - 4 * qpp + qppp) ** 2
    b2 = k1 * (q0 - 2 * qp + qpp) ** 2 + k2 * (q0 - qpp) ** 2
    b3 = k1 * (qm - 2 * q0 + qp) ** 2 + k2 * (qm - 4 * q0 + 3 * qp) ** 2
    return qi1, qi2, qi3, b1, b2, b3


def weno_5pts_right(qm, q0, qp, qpp, qppp):
    """``advection/weno.py:462-491`` (right-biased, for negative flow)."""
    eps = 1e-8
    qi1, qi2, qi3, b1, b2, b3 = _weno5_right_parts(qm, q0, qp, qpp, qppp)
    w1 = 0.1 / (b1 + eps) ** 2
    w2 = 0.6 / (b2 + eps) ** 2
    w3 = 0.3 / (b3 + eps) ** 2
    return (w1 * qi1 + w2 * qi2 + w3 * qi3) / (w1 + w2 + w3)


def wenoz_5pts(qmm, qm, q0, qp, qpp):
    """``advection/weno.py:379-410`` (tau5 = |b1 - b3|, eps 1e-16)."""
    eps = 1e-16
    qi1, qi2, qi3, b1, b2, b3 = _weno5_parts(qmm, qm, q0, qp, qpp)
    tau5 = np.abs(b1 - b3)
    w1 = 0.1 * (1 + tau5 / (b1 + eps))
    w2 = 0.6 * (1 + tau5 / (b2 + eps))
    w3 = 0.3 * (1 + tau5 / (b3 + eps))
    return (w1 * qi1 + w2 * qi2 + w3 * qi3) / (w1 + w2 + w3)


def wenoz_5pts_right(qm, q0, qp, qpp, qppp):
    """``advection/weno.py:494-522``."""
    eps = 1e-16
    qi1, qi2, qi3, b1, b2, b3 = _weno5_right_parts(qm, q0, qp, qpp, qppp)
    tau5 = np.abs(b1 - b3)
    w1 = 0.1 * (1 + tau5 / (b1 + eps))
    w2 = 0.6 * (1 + tau5 / (b2 + eps))
    w3 = 0.3 * (1 + tau5 / (b3 + eps))
    return (w1 * qi1 + w2 * qi2 + w3 * qi3) / (w1 + w2 + w3)


def minmod(r):
    """``advection/limiters.py:34-52``."""
    return np.maximum(0.0, np.minimum(1.0, r))


def van_leer(r):
    """``advection/limiters.py:55-74``."""
    ra = np.abs(r)
    return (r + ra) / (1.0 + ra)


def superbee(r):
    """``advection/limiters.py:77-98``."""
    return np.maximum(0.0, np.maximum(np.minimum(2.0 * r, 1.0), np.minimum(r, 2.0)))


def mc(r):
    """``advection/limiters.py:101-123``."""
    return np.maximum(0.0, np.minimum(np.minimum((1.0 + r) / 2.0, 2.0 * r), 2.0))


LIMITERS = {"minmod": minmod, "van_leer": van_leer, "superbee": superbee, "mc": mc}
TVD_EPS = 1e-8  # advection/reconstruction.py:43

# ---------------------------------------------------------------------------
# Face-flux reconstruction.  Written once along the last axis; the y
# direction is the x code on transposed arrays (the reference writes
# both out; ``advection/reconstruction.py:582-1102``).
# ---------------------------------------------------------------------------


def _face_x(h, u, method, limiter=None):
    """East-face flux on rows 1..Ny-2, faces i = 1..Nx-2; ghost ring zero.

    Restates ``Reconstruction2D.{naive,upwind1,upwind2,upwind3,weno3,
    wenoz3,weno5,wenoz5,tvd}_x`` (``advection/reconstruction.py:582-1050``).
    """
    H = h[1:-1]
    uf = u[1:-1, 1:-1]
    if method == "naive":
        face = 0.5 * (H[:, 1:-1] + H[:, 2:])
        return interior(face * uf, h.shape)
    if method == "upwind1":
        pos = H[:, 1:-1]
        neg = H[:, 2:]
    elif method in ("upwind2", "upwind3", "weno3", "wenoz3") or method == "tvd":
        if method == "upwind2":
            pos = 1.5 * H[:, 1:-1] - 0.5 * H[:, :-2]
            neg_int = 1.5 * H[:, 2:-1] - 0.5 * H[:, 3:]
        elif method == "upwind3":
            pos = -1.0 / 6.0 * H[:, :-2] + 5.0 / 6.0 * H[:, 1:-1] + 1.0 / 3.0 * H[:, 2:]
            neg_int = 1.0 / 3.0 * H[:, 1:-2] + 5.0 / 6.0 * H[:, 2:-1] - 1.0 / 6.0 * H[:, 3:]
        elif method == "weno3":
            pos = weno_3pts(H[:, :-2], H[:, 1:-1], H[:, 2:])
            neg_int = weno_3pts_right(H[:, 1:-2], H[:, 2:-1], H[:, 3:])
        elif method == "wenoz3":
            pos = wenoz_3pts(H[:, :-2], H[:, 1:-1], H[:, 2:])
            neg_int = wenoz_3pts_right(H[:, 1:-2], H[:, 2:-1], H[:, 3:])
        else:
            phi = LIMITERS[limiter]
            dfwd = H[:, 2:] - H[:, 1:-1]
            dbwd = H[:, 1:-1] - H[:, :-2]
            pos = H[:, 1:-1] + 0.5 * phi(dbwd / (dfwd + TVD_EPS)) * dfwd
            dneg = H[:, 1:-2] - H[:, 2:-1]
            dneg2 = H[:, 2:-1] - H[:, 3:]
            neg_int = H[:, 2:-1] + 0.5 * phi(dneg2 / (dneg + TVD_EPS)) * dneg
        neg = np.concatenate([neg_int, H[:, -1:]], axis=1)
    elif method in ("weno5", "wenoz5"):
        if method == "weno5":
            f5, f5r, f3, f3r = weno_5pts, weno_5pts_right, weno_3pts, weno_3pts_right
        else:
            f5, f5r, f3, f3r = wenoz_5pts, wenoz_5pts_right, wenoz_3pts, wenoz_3pts_right
        s = (H[:, :-4], H[:, 1:-3], H[:, 2:-2], H[:, 3:-1], H[:, 4:])
        pos = np.concatenate([f3(H[:, 0:1], H[:, 1:2], H[:, 2:3]), f5(*s),
                              f3(H[:, -3:-2], H[:, -2:-1], H[:, -1:])], axis=1)
        neg = np.concatenate([f5r(*s), f3r(H[:, -3:-2], H[:, -2:-1], H[:, -1:]),
                              H[:, -1:]], axis=1)
    else:
        raise ValueError(f"Unknown method: {method!r}")
    face = np.where(uf >= 0.0, pos, neg)
    return interior(face * uf, h.shape)


def face_flux_x(h, u, method, limiter=None):
    """East-face flux (see ``_face_x``)."""
    return _face_x(np.asarray(h, np.float64), np.asarray(u, np.float64), method, limiter)


def face_flux_y(h, v, method, limiter=None):
    """North-face flux: the x reconstruction on the transposed arrays."""
    return _face_x(np.asarray(h, np.float64).T, np.asarray(v, np.float64).T, method, limiter).T


TVD_NAMES = ("minmod", "van_leer", "superbee", "mc")
MASK_DISPATCHABLE = ("weno3", "weno5", "wenoz5") + TVD_NAMES  # advection/advection.py:40-43


def _tier_funcs(method):
    """``advection/advection.py:91-136``: stencil size -> (method, limiter)."""
    if method == "weno5":
        return {2: ("upwind1", None), 4: ("weno3", None), 6: ("weno5", None)}
    if method == "wenoz5":
        return {2: ("upwind1", None), 4: ("wenoz3", None), 6: ("wenoz5", None)}
    if method == "weno3":
        return {2: ("upwind1", None), 4: ("weno3", None)}
    if method in TVD_NAMES:
        return {2: ("upwind1", None), 4: ("tvd", method)}
    raise ValueError(f"Method {method!r} does not support mask-based stencil dispatch")


def upwind_flux(q, u, dim, funcs, hier):
    """``advection/flux.py:100-253``: per face, the tier of the upwind cell."""
    flux = face_flux_x if dim == 1 else face_flux_y
    fluxes = {s: flux(q, u, m, lim) for s, (m, lim) in funcs.items()}
    pos = u[1:-1, 1:-1] >= 0.0
    acc = np.zeros_like(q[1:-1, 1:-1], dtype=np.float64)
    for s in sorted(funcs):
        m = hier[s]
        shifted = m[1:-1, 2:] if dim == 1 else m[2:, 1:-1]
        fm = np.where(pos, m[1:-1, 1:-1], shifted)
        acc = acc + np.where(fm, fluxes[s][1:-1, 1:-1], 0.0)
    out = np.zeros(q.shape)
    out[1:-1, 1:-1] = acc
    return out


def advection_2d(h, u, v, dx, dy, method="upwind1", mask=None):
    """``Advection2D.__call__`` (``advection/advection.py:329-420``).

    Tendency -div(h u) written to [2:-2, 2:-2]; masked dispatchable methods
    route through ``upwind_flux`` with the (2, 4, 6) tier hierarchy built in
    ``__init__`` (``:310-327``); the output is multiplied by ``mask.h``.
    """
    h = np.asarray(h, np.float64)
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    if mask is not None and method in MASK_DISPATCHABLE:
        funcs = _tier_funcs(method)
        sizes = tuple(sorted(funcs))
        hx = narrow_mask_hierarchy(mask.get_adaptive_masks("x", (2, 4, 6)), sizes)
        hy = narrow_mask_hierarchy(mask.get_adaptive_masks("y", (2, 4, 6)), sizes)
        fe = upwind_flux(h, u, 1, funcs, hx)
        fn = upwind_flux(h, v, 0, funcs, hy)
    else:
        if method in TVD_NAMES:
            fe = face_flux_x(h, u, "tvd", method)
            fn = face_flux_y(h, v, "tvd", method)
        elif method in ("naive", "upwind1", "upwind2", "upwind3", "weno3", "weno5", "wenoz5"):
            fe = face_flux_x(h, u, method)
            fn = face_flux_y(h, v, method)
        else:
            raise ValueError(f"Unknown method: {method!r}")
    out = np.zeros(h.shape)
    out[2:-2, 2:-2] = -((fe[2:-2, 2:-2] - fe[2:-2, 1:-3]) / dx + (fn[2:-2, 2:-2] - fn[1:-3, 2:-2]) / dy)
    if mask is not None:
        out = out * mask.h
    return out


# ---------------------------------------------------------------------------
# Arakawa Jacobian
# ---------------------------------------------------------------------------


def arakawa_jacobian(f, g, dx, dy):
    """``operators/jacobian.py:22-131``: (Jpp + Jpx + Jxp) / (12 dx dy).

    Output is interior-only, shape ``(..., Ny-2, Nx-2)``.
    """
    f = np.asarray(f, np.float64)
    g = np.asarray(g, np.float64)
    E = (Ellipsis, slice(1, -1), slice(2, None))
    W = (Ellipsis, slice(1, -1), slice(None, -2))
    N = (Ellipsis, slice(2, None), slice(1, -1))
    S = (Ellipsis, slice(None, -2), slice(1, -1))
    NE = (Ellipsis, slice(2, None), slice(2, None))
    NW = (Ellipsis, slice(2, None), slice(None, -2))
    SE = (Ellipsis, slice(None, -2), slice(2, None))
    SW = (Ellipsis, slice(None, -2), slice(None, -2))
    jpp = (f[E] - f[W]) * (g[N] - g[S]) - (f[N] - f[S]) * (g[E] - g[W])
    jpx = (f[E] * (g[NE] - g[SE]) - f[W] * (g[NW] - g[SW])
           - f[N] * (g[NE] - g[NW]) + f[S] * (g[SE] - g[SW]))
    jxp = (g[N] * (f[NE] - f[NW]) - g[S] * (f[SE] - f[SW])
           - g[E] * (f[NE] - f[SE]) + g[W] * (f[NW] - f[SW]))
    return (jpp + jpx + jxp) / (12.0 * dx * dy)


# ---------------------------------------------------------------------------
# Diffusion
# ---------------------------------------------------------------------------


def diffusion_fluxes(h, kappa, dx, dy, mask=None):
    """``diffusion/diffusion.py:179-206``: faces strictly between interior cells."""
    h = np.asarray(h, np.float64)
    k = np.asarray(kappa, np.float64)
    kx = k[1:-1, 1:-2] if k.ndim >= 2 else k
    ky = k[1:-2, 1:-1] if k.ndim >= 2 else k
    fx = np.zeros(h.shape)
    fx[1:-1, 1:-2] = kx * (h[1:-1, 2:-1] - h[1:-1, 1:-2]) / dx
    fy = np.zeros(h.shape)
    fy[1:-2, 1:-1] = ky * (h[2:-1, 1:-1] - h[1:-2, 1:-1]) / dy
    if mask is not None:
        fx = fx * mask.u
        fy = fy * mask.v
    return fx, fy


def diffusion_2d(h, kappa, dx, dy, mask=None):
    """``diffusion/diffusion.py:120-176``: no-flux walls, masked fluxes."""
    fx, fy = diffusion_fluxes(h, kappa, dx, dy, mask)
    out = interior((fx[1:-1, 1:-1] - fx[1:-1, :-2]) / dx + (fy[1:-1, 1:-1] - fy[:-2, 1:-1]) / dy,
                   h.shape)
    if mask is not None:
        out = out * mask.h
    return out


def biharmonic_2d(h, kappa, dx, dy, mask=None):
    """``diffusion/diffusion.py:532-574``: -kappa * D(D(h)), D unmasked."""
    lap1 = diffusion_2d(h, 1.0, dx, dy)
    out = -kappa * diffusion_2d(lap1, 1.0, dx, dy)
    if mask is not None:
        out = out * mask.h
    return out


# ---------------------------------------------------------------------------
# Elliptic operators and multigrid
# ---------------------------------------------------------------------------


def masked_laplacian(psi, mask, dx, dy, lambda_=0.0):
    """``solvers/iterative.py:144-191``: Dirichlet outside the mask, periodic roll."""
    m = np.asarray(mask, np.float64)
    p = np.asarray(psi, np.float64) * m
    lap = (np.roll(p, 1, axis=1) + np.roll(p, -1, axis=1) - 2.0 * p) / dx**2 + \
          (np.roll(p, 1, axis=0) + np.roll(p, -1, axis=0) - 2.0 * p) / dy**2
    return (lap - lambda_ * p) * m


def compute_n_levels(ny, nx, min_coarse=8):
    """``solvers/multigrid.py:50-76``."""
    n = 1
    while ny % 2 == 0 and nx % 2 == 0 and ny // 2 >= min_coarse and nx // 2 >= min_coarse:
        ny //= 2
        nx //= 2
        n += 1
    return n


def restrict_mask(mask, threshold=0.5):
    """``solvers/multigrid.py:79-101``: coarse wet iff >= 2 of 4 fine wet."""
    avg = (mask[::2, ::2] + mask[1::2, ::2] + mask[::2, 1::2] + mask[1::2, 1::2]) / 4.0
    return (avg >= threshold).astype(np.float64)


def interpolate_coeff_to_faces(coeff, mask):
    """``solvers/multigrid.py:104-179``: arithmetic face mean x face mask."""
    mcx = np.zeros_like(mask)
    mcx[:, :-1] = mask[:, :-1] * mask[:, 1:]
    mcy = np.zeros_like(mask)
    mcy[:-1, :] = mask[:-1, :] * mask[1:, :]
    cx = np.zeros_like(coeff)
    cx[:, :-1] = 0.5 * (coeff[:, :-1] + coeff[:, 1:])
    cy = np.zeros_like(coeff)
    cy[:-1, :] = 0.5 * (coeff[:-1, :] + coeff[1:, :])
    return cx * mcx, cy * mcy


def restrict_coeff(coeff, mf, mc_):
    """``solvers/multigrid.py:182-218``."""
    cm = coeff * mf
    s4 = cm[::2, ::2] + cm[1::2, ::2] + cm[::2, 1::2] + cm[1::2, 1::2]
    cnt = mf[::2, ::2] + mf[1::2, ::2] + mf[::2, 1::2] + mf[1::2, 1::2]
    return (s4 / np.maximum(cnt, 1.0)) * mc_


def compute_inv_diagonal(cx, cy, mask, dx, dy, lambda_):
    """``solvers/multigrid.py:221-279``."""
    cxw = np.zeros_like(cx)
    cxw[:, 1:] = cx[:, :-1]
    cys = np.zeros_like(cy)
    cys[1:, :] = cy[:-1, :]
    diag = (-(cx + cxw) / dx**2 - (cy + cys) / dy**2 - lambda_) * mask
    safe = np.where(np.abs(diag) > 1e-30, diag, 1.0)
    return np.where(np.abs(diag) > 1e-30, 1.0 / safe, 0.0) * mask


class Level:
    """One multigrid level (``solvers/multigrid.py:283-322``).

    ``kind='noflux'`` is the reference operator (face coefficients zero
    across land faces and at the array edge).  ``kind='dirichlet'`` is the
    constant-coefficient ``masked_laplacian`` operator with zero padding at
    the array edge; it has no reference counterpart in the multigrid module
    and is this build's hierarchy for the Dirichlet QG inversion.
    """

    def __init__(self, mask, cx, cy, dx, dy, lambda_, inv_diag, kind, wall_g=0.0):
        self.mask, self.cx, self.cy = mask, cx, cy
        self.dx, self.dy, self.lambda_ = dx, dy, lambda_
        self.inv_diag, self.kind = inv_diag, kind
        self.wall_g = wall_g


def wall_counts(m):
    """Numbers of dry (or outside) x and y neighbours of every cell."""
    mp = np.pad(m, 1)
    nlx = 2.0 - mp[1:-1, 2:] - mp[1:-1, :-2]
    nly = 2.0 - mp[2:, 1:-1] - mp[:-2, 1:-1]
    return nlx, nly


def wall_g(level):
    """Coarse-level wall correction of the Dirichlet hierarchy (no reference
    counterpart).  The fine operator's zero ghost puts psi = 0 one fine cell
    beyond the last wet cell centre; on level l (cells 2^l times larger) the
    ghost value that extrapolates a linear profile to that same point is
    g u_0 with g = (1 - 2^l) / (1 + 2^l) (0 on level 0)."""
    r = 2.0 ** level
    return (1.0 - r) / (1.0 + r)


def build_multigrid(mask, dx, dy, lambda_=0.0, coeff=None, n_levels=None, kind="noflux",
                    wall=False):
    """``build_multigrid_solver`` (``solvers/multigrid.py:905-1062``).  ``wall``
    (Dirichlet hierarchies only, no reference counterpart): coarse levels take
    the wall correction of ``wall_g`` on cells beside land or the edge."""
    m = np.asarray(mask, np.float64)
    ny, nx = m.shape
    if n_levels is None:
        n_levels = compute_n_levels(ny, nx)
    f = 2 ** (n_levels - 1)
    if ny % f or nx % f:
        raise ValueError(f"Grid shape ({ny}, {nx}) is not divisible by 2^(n_levels-1) = {f}.")
    c = np.ones_like(m) if coeff is None else np.asarray(coeff, np.float64)
    c = c * m
    levels = []
    cur_m, cur_c, cdx, cdy = m, c, float(dx), float(dy)
    for lev in range(n_levels):
        if kind == "noflux":
            cx, cy = interpolate_coeff_to_faces(cur_c, cur_m)
            inv = compute_inv_diagonal(cx, cy, cur_m, cdx, cdy, lambda_)
        else:
            cx = cy = None
            g = wall_g(lev) if wall else 0.0
            nlx, nly = wall_counts(cur_m)
            d = -2.0 / cdx**2 - 2.0 / cdy**2 - lambda_ + g * (nlx / cdx**2 + nly / cdy**2)
            inv = cur_m / d
        levels.append(Level(cur_m, cx, cy, cdx, cdy, float(lambda_), inv, kind,
                            g if kind != "noflux" else 0.0))
        if lev < n_levels - 1:
            nm = restrict_mask(cur_m)
            cur_c = restrict_coeff(cur_c, cur_m, nm)
            cur_m = nm
            cdx *= 2.0
            cdy *= 2.0
    return levels


def apply_operator(u, lev):
    """``solvers/multigrid.py:327-405`` (noflux) / masked_laplacian with zero pad."""
    m = lev.mask
    um = u * m
    ue = np.zeros_like(um)
    ue[:, :-1] = um[:, 1:]
    uw = np.zeros_like(um)
    uw[:, 1:] = um[:, :-1]
    un = np.zeros_like(um)
    un[:-1, :] = um[1:, :]
    us = np.zeros_like(um)
    us[1:, :] = um[:-1, :]
    if lev.kind == "noflux":
        cx, cy = lev.cx, lev.cy
        cxw = np.zeros_like(cx)
        cxw[:, 1:] = cx[:, :-1]
        cys = np.zeros_like(cy)
        cys[1:, :] = cy[:-1, :]
        lap = (cx * (ue - um) - cxw * (um - uw)) / lev.dx**2 + \
              (cy * (un - um) - cys * (um - us)) / lev.dy**2
    else:
        lap = (ue + uw - 2.0 * um) / lev.dx**2 + (un + us - 2.0 * um) / lev.dy**2
        if lev.wall_g:
            nlx, nly = wall_counts(m)
            lap = lap + lev.wall_g * (nlx / lev.dx**2 + nly / lev.dy**2) * um
    return (lap - lev.lambda_ * um) * m


def weighted_jacobi(u, rhs, lev, omega, n):
    """``solvers/multigrid.py:408-461``; ``omega`` may also be a sequence of
    per-sweep weights (sweep k takes omega[k mod len])."""
    ws = list(omega) if np.ndim(omega) else [omega]
    for k in range(n):
        u = (u + ws[k % len(ws)] * lev.inv_diag * (rhs - apply_operator(u, lev))) * lev.mask
    return u


def chebyshev_weights(lo, m, hi=2.0):
    """Per-sweep weights of an m-sweep polynomial (Chebyshev) smoother on
    [lo, hi] of D^-1 A (no reference counterpart): the reciprocals of the
    roots of the degree-m Chebyshev polynomial of that interval, ascending."""
    k = np.arange(m)
    roots = 0.5 * (hi + lo) + 0.5 * (hi - lo) * np.cos(np.pi * (2 * k + 1) / (2 * m))
    return list(1.0 / roots)


def restrict(v, mf, mc_):
    """``solvers/multigrid.py:464-507``: mask-weighted 2x2 mean."""
    vm = v * mf
    s4 = vm[::2, ::2] + vm[1::2, ::2] + vm[::2, 1::2] + vm[1::2, 1::2]
    cnt = mf[::2, ::2] + mf[1::2, ::2] + mf[::2, 1::2] + mf[1::2, 1::2]
    return (s4 / np.maximum(cnt, 1.0)) * mc_


def prolongate(vc, mc_, mf):
    """``solvers/multigrid.py:510-603``: 9/3/3/1 weights, mask renormalised."""
    nyc, nxc = vc.shape
    vp = np.pad(vc * mc_, 1)
    mp = np.pad(mc_, 1)
    out = np.zeros((2 * nyc, 2 * nxc))
    for dj in (0, 1):
        for di in (0, 1):
            sj, si = 2 * dj - 1, 2 * di - 1
            vx = vp[1:-1, 1 + si:nxc + 1 + si]
            mx = mp[1:-1, 1 + si:nxc + 1 + si]
            vy = vp[1 + sj:nyc + 1 + sj, 1:-1]
            my = mp[1 + sj:nyc + 1 + sj, 1:-1]
            vd = vp[1 + sj:nyc + 1 + sj, 1 + si:nxc + 1 + si]
            md = mp[1 + sj:nyc + 1 + sj, 1 + si:nxc + 1 + si]
            w = 9 * vp[1:-1, 1:-1] + 3 * vx + 3 * vy + vd
            d = 9 * mp[1:-1, 1:-1] + 3 * mx + 3 * my + md
            out[dj::2, di::2] = w / np.maximum(d, 1.0)
    return out * mf


def restrict_dirichlet(v, mf, mc_):
    """The Dirichlet hierarchy's restriction (no reference counterpart): the
    2x2 sum of the wet children over 4 (dry children hold zero, as psi on
    land), not renormalised by the number of wet children."""
    vm = v * mf
    return 0.25 * (vm[::2, ::2] + vm[1::2, ::2] + vm[::2, 1::2] + vm[1::2, 1::2]) * mc_


def prolongate_dirichlet(vc, mc_, mf):
    """The Dirichlet hierarchy's prolongation (no reference counterpart):
    the 9/3/3/1 weights of ``prolongate`` over 16 everywhere, land and the
    space beyond the walls contributing zero (psi = 0 there), instead of
    renormalising by the wet weights, which would extrapolate the coarse
    correction to the wall."""
    nyc, nxc = vc.shape
    vp = np.pad(vc * mc_, 1)
    out = np.zeros((2 * nyc, 2 * nxc))
    for dj in (0, 1):
        for di in (0, 1):
            sj, si = 2 * dj - 1, 2 * di - 1
            w = (9 * vp[1:-1, 1:-1] + 3 * vp[1:-1, 1 + si:nxc + 1 + si]
                 + 3 * vp[1 + sj:nyc + 1 + sj, 1:-1] + vp[1 + sj:nyc + 1 + sj, 1 + si:nxc + 1 + si])
            out[dj::2, di::2] = w / 16.0
    return out * mf


def v_cycle(levels, u, rhs, idx=0, n_pre=6, n_post=6, n_coarse=50, omega=0.95, cheb=None):
    """``MultigridSolver.v_cycle`` (``solvers/multigrid.py:808-883``); the
    Dirichlet hierarchy transfers with ``restrict_dirichlet`` and
    ``prolongate_dirichlet``.  ``cheb`` (no reference counterpart): the lower
    end of the Chebyshev smoothing interval [cheb, 2]; pre-smoothing and the
    bottom take the weights ascending, post-smoothing descending."""
    lev = levels[idx]
    w_pre = omega if cheb is None else chebyshev_weights(cheb, n_pre)
    w_post = omega if cheb is None else chebyshev_weights(cheb, n_post)[::-1]
    w_bot = omega if cheb is None else chebyshev_weights(cheb, n_coarse)
    if idx == len(levels) - 1:
        return weighted_jacobi(u, rhs, lev, w_bot, n_coarse)
    u = weighted_jacobi(u, rhs, lev, w_pre, n_pre)
    r = (rhs - apply_operator(u, lev)) * lev.mask
    cl = levels[idx + 1]
    Rs = restrict_dirichlet if lev.kind == "dirichlet" else restrict
    rc = Rs(r, lev.mask, cl.mask)
    ec = v_cycle(levels, np.zeros_like(rc), rc, idx + 1, n_pre, n_post, n_coarse, omega, cheb)
    P = prolongate_dirichlet if lev.kind == "dirichlet" else prolongate
    u = (u + P(ec, cl.mask, lev.mask)) * lev.mask
    return weighted_jacobi(u, rhs, lev, w_post, n_post)


def mg_solve(levels, rhs, n_cycles=5, **kw):
    """``MultigridSolver._run_vcycles`` (``solvers/multigrid.py:885-902``)."""
    u = np.zeros_like(rhs)
    for _ in range(n_cycles):
        u = v_cycle(levels, u, rhs, **kw)
    return u


def pcg(matvec, b, x0=None, precond=None, rtol=1e-6, atol=1e-6, max_steps=500,
        flexible=True):
    """Preconditioned CG for the negative-definite operator ``A``.

    The reference (``solvers/iterative.py:48-136``) hands the NSD operator
    and the negated preconditioner to lineax 0.1.0's CG; lineax's iterate
    sequence and stopping rule are not restated here (parity unpinned).
    This build's rule: stop when ||b - A x||_2 <= max(rtol ||b||_2, atol).
    ``flexible=True`` uses the Polak-Ribiere beta, which tolerates the
    non-symmetric V-cycle; the result is pinned by direct sparse solves.
    Returns ``(x, iterations, residual_norm, converged)``.
    """
    b = np.asarray(b, np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64)
    M = (lambda r: r) if precond is None else precond
    # A and M are both negative definite; CG on (A, M) gives the same
    # iterates as CG on the positive pair (-A, -M).
    r = b - matvec(x)
    bn = np.linalg.norm(b)
    tol = max(rtol * bn, atol)
    rn = np.linalg.norm(r)
    if rn <= tol:
        return x, 0, rn, True
    z = M(r)
    p = z.copy()
    rz = np.vdot(r, z)
    for k in range(1, max_steps + 1):
        q = matvec(p)
        alpha = rz / np.vdot(p, q)
        x = x + alpha * p
        r_old = r
        r = r - alpha * q
        rn = np.linalg.norm(r)
        if rn <= tol:
            return x, k, np.linalg.norm(b - matvec(x)), True
        z = M(r)
        rz_new = np.vdot(r, z)
        beta = (np.vdot(z, r - r_old) if flexible else rz_new) / rz
        rz = rz_new
        p = z + beta * p
    return x, max_steps, np.linalg.norm(b - matvec(x)), False


def operator_matrix(apply, shape):
    """Dense-by-columns sparse assembly of a linear map (brute force)."""
    import scipy.sparse as sp

    n = shape[0] * shape[1]
    cols = []
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        cols.append(sp.csc_matrix(apply(e.reshape(shape)).reshape(n, 1)))
    return sp.hstack(cols).tocsr()


def level_matrix(lev):
    """Sparse (ny*nx)^2 matrix of ``apply_operator`` at one level, assembled
    directly from the stencil coefficients (vectorised)."""
    import scipy.sparse as sp

    m = lev.mask
    ny, nx = m.shape
    n = ny * nx
    idx = np.arange(n).reshape(ny, nx)
    ix2, iy2 = 1.0 / lev.dx**2, 1.0 / lev.dy**2
    rows, cols, vals = [], [], []

    def add(r, c, v):
        rows.append(r.ravel())
        cols.append(c.ravel())
        vals.append(v.ravel())

    if lev.kind == "noflux":
        cx, cy = lev.cx, lev.cy
        cxw = np.zeros_like(cx)
        cxw[:, 1:] = cx[:, :-1]
        cys = np.zeros_like(cy)
        cys[1:, :] = cy[:-1, :]
        diag = (-(cx + cxw) * ix2 - (cy + cys) * iy2 - lev.lambda_) * m * m
        add(idx, idx, diag)
        add(idx[:, :-1], idx[:, 1:], (cx[:, :-1] * ix2) * m[:, :-1] * m[:, 1:])
        add(idx[:, 1:], idx[:, :-1], (cxw[:, 1:] * ix2) * m[:, 1:] * m[:, :-1])
        add(idx[:-1, :], idx[1:, :], (cy[:-1, :] * iy2) * m[:-1, :] * m[1:, :])
        add(idx[1:, :], idx[:-1, :], (cys[1:, :] * iy2) * m[1:, :] * m[:-1, :])
    else:
        nlx, nly = wall_counts(m)
        diag = (-2.0 * ix2 - 2.0 * iy2 - lev.lambda_
                + lev.wall_g * (nlx * ix2 + nly * iy2)) * m * m
        add(idx, idx, diag)
        add(idx[:, :-1], idx[:, 1:], ix2 * m[:, :-1] * m[:, 1:])
        add(idx[:, 1:], idx[:, :-1], ix2 * m[:, 1:] * m[:, :-1])
        add(idx[:-1, :], idx[1:, :], iy2 * m[:-1, :] * m[1:, :])
        add(idx[1:, :], idx[:-1, :], iy2 * m[1:, :] * m[:-1, :])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


def dirichlet_helmholtz_matrix(mask, dx, dy, lambda_):
    """Sparse (wet x wet) matrix of the zero-padded masked Dirichlet operator."""
    import scipy.sparse as sp

    m = np.asarray(mask) > 0
    ny, nx = m.shape
    idx = -np.ones((ny, nx), dtype=np.int64)
    idx[m] = np.arange(m.sum())
    rows, cols, vals = [], [], []
    diag = -2.0 / dx**2 - 2.0 / dy**2 - lambda_
    jj, ii = np.nonzero(m)
    k = idx[jj, ii]
    rows.append(k)
    cols.append(k)
    vals.append(np.full(k.size, diag))
    for dj, di, w in ((0, 1, 1 / dx**2), (0, -1, 1 / dx**2), (1, 0, 1 / dy**2), (-1, 0, 1 / dy**2)):
        j2, i2 = jj + dj, ii + di
        ok = (j2 >= 0) & (j2 < ny) & (i2 >= 0) & (i2 < nx)
        ok[ok] = m[j2[ok], i2[ok]]
        rows.append(k[ok])
        cols.append(idx[j2[ok], i2[ok]])
        vals.append(np.full(ok.sum(), w))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(m.sum(), m.sum()))
    return A, idx


def dirichlet_helmholtz_direct(rhs, mask, dx, dy, lambda_):
    """Exact solve of the zero-padded masked Dirichlet Helmholtz problem."""
    import scipy.sparse.linalg as spla

    m = np.asarray(mask) > 0
    A, _ = dirichlet_helmholtz_matrix(m, dx, dy, lambda_)
    x = np.zeros(m.shape)
    x[m] = spla.spsolve(A.tocsc(), np.asarray(rhs, np.float64)[m])
    return x


# ---------------------------------------------------------------------------
# Time stepping, ghost fill, vertical modes
# ---------------------------------------------------------------------------


def euler_step(state, rhs_fn, dt):
    """``timestepping/explicit_rk.py:22-40``."""
    return state + dt * rhs_fn(state)


def heun_step(state, rhs_fn, dt):
    """``timestepping/explicit_rk.py:43-71`` (Heun / RK2)."""
    k1 = rhs_fn(state)
    k2 = rhs_fn(state + dt * k1)
    return state + 0.5 * dt * (k1 + k2)


def rk4_step(state, rhs_fn, dt):
    """``timestepping/explicit_rk.py:124-168`` (classical RK4)."""
    k1 = rhs_fn(state)
    k2 = rhs_fn(state + 0.5 * dt * k1)
    k3 = rhs_fn(state + 0.5 * dt * k2)
    k4 = rhs_fn(state + dt * k3)
    return state + dt / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4)


def rk3_ssp_step(state, rhs_fn, dt):
    """``timestepping/explicit_rk.py:74-121`` (Shu-Osher SSP-RK3)."""
    k1 = rhs_fn(state)
    y1 = state + dt * k1
    k2 = rhs_fn(y1)
    y2 = 0.75 * state + 0.25 * y1 + 0.25 * dt * k2
    k3 = rhs_fn(y2)
    return (1.0 / 3.0) * state + (2.0 / 3.0) * y2 + (2.0 / 3.0) * dt * k3


def enforce_periodic(field, ng=1):
    """Periodic ghost fill of width ``ng`` (``boundary/boundary.py:52-78``)."""
    f = np.array(field, np.float64)
    f[:ng, :] = f[-2 * ng:-ng, :]
    f[-ng:, :] = f[ng:2 * ng, :]
    f[:, :ng] = f[:, -2 * ng:-ng]
    f[:, -ng:] = f[:, ng:2 * ng]
    return f


def build_coupling_matrix(H, g_prime):
    """``vertical/vertical_modes.py:50-121``."""
    H = np.asarray(H, np.float64)
    g = np.asarray(g_prime, np.float64)
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


def decompose_vertical_modes(A, f0):
    """``vertical/vertical_modes.py:124-190``.

    ``jnp.linalg.eigh`` symmetrises its input, so the oracle eigendecomposes
    ``(A + A^T)/2``; eigenvector signs are arbitrary (compare up to sign).
    """
    As = 0.5 * (np.asarray(A, np.float64) + np.asarray(A, np.float64).T)
    lam, R = np.linalg.eigh(As)
    pos = lam > 0
    radii = np.where(pos, 1.0 / (abs(f0) * np.sqrt(np.where(pos, lam, 1.0))), np.inf)
    return radii, R.T, R


def layer_to_mode(field, Cl2m):
    """``vertical/vertical_modes.py:193-227``."""
    nl = field.shape[0]
    return (Cl2m @ field.reshape(nl, -1)).reshape(field.shape)


def mode_to_layer(field, Cm2l):
    """``vertical/vertical_modes.py:230-263``."""
    nl = field.shape[0]
    return (Cm2l @ field.reshape(nl, -1)).reshape(field.shape)


# ---------------------------------------------------------------------------
# QG 1.5-layer step (build-defined composition; nearest reference:
# scripts/qg_1p5_layer.py:420-497, which uses upwind1 + DST + Heun instead)
# ---------------------------------------------------------------------------


class QGParams:
    """Physical and numerical parameters of the 1.5-layer QG step.

    dq/dt = -J(psi, q + beta (y - y0)) + F - nu4 D(D(q)),
    (lap - lambda) psi = q with psi = 0 outside the wet mask and beyond the
    basin walls; J is ``arakawa_jacobian``, D the no-flux ``Diffusion2D``
    (``BiharmonicDiffusion2D``), stepping by ``rk3_ssp_step``.
    """

    def __init__(self, nx, ny, Lx, Ly, f0=9.375e-5, beta=1.754e-11, Ld=4e4, nu4=1e9,
                 F0=2e-12, dt=1200.0):
        self.nx, self.ny, self.Lx, self.Ly = nx, ny, Lx, Ly
        self.dx, self.dy = Lx / nx, Ly / ny
        self.f0, self.beta, self.Ld, self.nu4, self.F0, self.dt = f0, beta, Ld, nu4, F0, dt
        self.lambda_ = 1.0 / Ld**2

    def y_centres(self):
        """Cell-centre y of rows 0..ny+1 (row 0 is the south ghost)."""
        return (np.arange(self.ny + 2) - 0.5) * self.dy

    def forcing(self):
        """Double-gyre wind-curl ``-F0 sin(2 pi y / Ly)`` (``qg_1p5_layer.py:363``)."""
        y = self.y_centres()
        F = np.zeros((self.ny + 2, self.nx + 2))
        F[1:-1, 1:-1] = (-self.F0 * np.sin(2.0 * np.pi * y[1:-1] / self.Ly))[:, None]
        return F

    def beta_y(self):
        y = self.y_centres()
        return np.broadcast_to((self.beta * (y - 0.5 * self.Ly))[:, None],
                               (self.ny + 2, self.nx + 2))


def qg_rhs(q, psi, p, F, wet):
    """QG tendency from q and the already-inverted psi (ghost rings zero)."""
    J = arakawa_jacobian(psi, q + p.beta_y(), p.dx, p.dy)
    bih = biharmonic_2d(q, p.nu4, p.dx, p.dy)
    out = np.zeros_like(q)
    out[1:-1, 1:-1] = (-J + F[1:-1, 1:-1] + bih[1:-1, 1:-1]) * wet
    return out


def qg_invert(q, p, wet, solver=None):
    """psi with (lap - lambda) psi = q on the wet interior, psi = 0 elsewhere."""
    rhs = q[1:-1, 1:-1] * wet
    if solver is None:
        psi_i = dirichlet_helmholtz_direct(rhs, wet, p.dx, p.dy, p.lambda_)
    else:
        psi_i = solver(rhs)
    psi = np.zeros_like(q)
    psi[1:-1, 1:-1] = psi_i * wet
    return psi


def qg_step(q, p, F, wet, solver=None):
    """One ``rk3_ssp_step`` of the QG model; returns the new q."""
    def rhs(qs):
        return qg_rhs(qs, qg_invert(qs, p, wet, solver), p, F, wet)
    return rk3_ssp_step(q, rhs, p.dt)


# ---------------------------------------------------------------------------
# The reference's 1.5-layer QG example (scripts/qg_1p5_layer.py): upwind1
# advection, geostrophic velocities through X points, DST Helmholtz
# inversion (here the exact sparse solve of the same Dirichlet operator),
# Heun stepping.
# ---------------------------------------------------------------------------


class QGScriptParams:
    """``QuasiGeostrophicConfig`` defaults (``scripts/qg_1p5_layer.py:280-300``)."""

    def __init__(self, nx=64, ny=64, Lx=5.12e6, Ly=5.12e6, beta=1.754e-11, rossby_radius=4.0e4,
                 drag=5.0e-8, viscosity=5.0e4, wind_curl_forcing=2.0e-12, dt=4000.0):
        self.nx, self.ny, self.Lx, self.Ly = nx, ny, Lx, Ly
        self.dx, self.dy = Lx / nx, Ly / ny
        self.beta, self.drag, self.viscosity, self.dt = beta, drag, viscosity, dt
        self.lambda_ = 1.0 / rossby_radius**2
        self.F0 = wind_curl_forcing

    def fields(self):
        """q0 and wind curl with zero ghost rings (``:332-371``)."""
        x = (np.arange(self.nx) + 0.5) * self.dx
        y = (np.arange(self.ny) + 0.5) * self.dy
        X, Y = np.meshgrid(x, y)
        q0 = np.pad(5.0e-9 * np.sin(2 * np.pi * X / self.Lx) * np.sin(np.pi * Y / self.Ly), 1)
        wind = np.pad(-self.F0 * np.sin(2 * np.pi * Y / self.Ly), 1)
        return q0, wind


def qg_script_invert(q, p):
    """``invert_streamfunction`` (``:432-452``)."""
    psi = np.zeros_like(q)
    psi[1:-1, 1:-1] = dirichlet_helmholtz_direct(q[1:-1, 1:-1], np.ones((p.ny, p.nx)), p.dx,
                                                 p.dy, p.lambda_)
    return psi


def qg_script_velocity(psi, p):
    """``geostrophic_velocity_from_streamfunction`` (``:115-145``)."""
    px = T_to_X(psi)
    return -diff_y_X_to_U(px, p.dy), diff_x_X_to_V(px, p.dx)


def qg_script_tendency(q, p, wind):
    """``tendency`` (``:454-478``): (q_rhs, psi, u, v)."""
    psi = qg_script_invert(q, p)
    u, v = qg_script_velocity(psi, p)
    q_rhs = advection_2d(q, u, v, p.dx, p.dy, "upwind1")
    q_rhs[1:-1, 1:-1] += -p.beta * V_to_T(v)[1:-1, 1:-1]
    q_rhs = q_rhs + wind - p.drag * laplacian(psi, p.dx, p.dy) + \
        p.viscosity * laplacian(q, p.dx, p.dy)
    return q_rhs, psi, u, v


def qg_script_step(q, p, wind):
    """``step`` (``:484-497``): Heun with the wall condition before each
    tendency and after the step; returns (q, psi, u, v)."""
    bc = lambda f: np.pad(f[1:-1, 1:-1], 1)  # noqa: E731  (pad_interior, zero ring)
    qn = bc(heun_step(q, lambda s: qg_script_tendency(bc(s), p, wind)[0], p.dt))
    psi = qg_script_invert(qn, p)
    u, v = qg_script_velocity(psi, p)
    return qn, psi, u, v


# ---------------------------------------------------------------------------
# Multi-layer QG with a passive tracer (build-defined composition of the
# reference's vertical-mode API, vertical/vertical_modes.py:50-263, with the
# 1.5-layer step above and Advection2D; no single reference function)
# ---------------------------------------------------------------------------


class MultiLayerParams(QGParams):
    """Layer stack and modes of the multi-layer QG step.

    Layer l obeys dq_l/dt = -J(psi_l, q_l + beta (y - y0)) + F delta_{l0}
    - nu4 D(D(q_l)), with the wind forcing on the top layer only.  The
    streamfunction comes from the vertical modes: ``radii, Cl2m, Cm2l =
    decompose_vertical_modes(build_coupling_matrix(H, g'), f0)``; mode k
    solves (lap - 1/radii_k^2) psi_k = (Cl2m q)_k (psi = 0 on land and the
    walls; an infinite radius gives a Poisson solve) and psi = Cm2l psi_modes.
    The tracer c is advected by the top layer's geostrophic face velocities
    (``scripts/qg_1p5_layer.py:115-145``) with masked Advection2D.
    """

    def __init__(self, nx, ny, Lx, Ly, H, g_prime, f0=9.375e-5, beta=1.754e-11, nu4=1e9,
                 F0=2e-12, dt=900.0, tracer_method="weno5"):
        super().__init__(nx, ny, Lx, Ly, f0=f0, beta=beta, Ld=1.0, nu4=nu4, F0=F0, dt=dt)
        self.H = np.asarray(H, np.float64)
        self.g_prime = np.asarray(g_prime, np.float64)
        A = build_coupling_matrix(self.H, self.g_prime)
        self.radii, self.Cl2m, self.Cm2l = decompose_vertical_modes(A, f0)
        self.mode_lambda = np.where(np.isfinite(self.radii), 1.0 / self.radii**2, 0.0)
        self.tracer_method = tracer_method

    @property
    def nl(self):
        return self.H.size


def multilayer_invert(q, p, wet):
    """Layer streamfunctions (nl, ny+2, nx+2) from layer PV through the
    vertical modes, one direct Dirichlet solve per mode."""
    qm = layer_to_mode(q, p.Cl2m)
    pm = np.zeros_like(q)
    for k in range(p.nl):
        pm[k, 1:-1, 1:-1] = dirichlet_helmholtz_direct(qm[k, 1:-1, 1:-1] * wet, wet, p.dx, p.dy,
                                                       p.mode_lambda[k]) * wet
    return mode_to_layer(pm, p.Cm2l)


def multilayer_rhs(q, psi, p, F, wet):
    """Per-layer QG tendency; forcing on layer 0 (the top) only."""
    zero = np.zeros_like(F)
    return np.stack([qg_rhs(q[l], psi[l], p, F if l == 0 else zero, wet) for l in range(p.nl)])


def tracer_velocity(psi_top, p, mask=None):
    """Top-layer face velocities (``scripts/qg_1p5_layer.py:115-145``, with
    the masked operators when a mask is given)."""
    px = T_to_X(psi_top, mask)
    return -diff_y_X_to_U(px, p.dy, mask), diff_x_X_to_V(px, p.dx, mask)


def multilayer_step(q, c, p, F, wet, mask=None):
    """One ``rk3_ssp_step`` of the layer PV ``q`` (nl, ny+2, nx+2) and the
    tracer ``c`` (ny+2, nx+2) together; returns (q, c, psi of the last
    stage)."""
    last = {}

    def rhs(state):
        qs, cs = state
        psi = multilayer_invert(qs, p, wet)
        u, v = tracer_velocity(psi[0], p, mask)
        last["psi"] = psi
        return (multilayer_rhs(qs, psi, p, F, wet),
                advection_2d(cs, u, v, p.dx, p.dy, p.tracer_method, mask))

    def comb(a, x, b, y, cdt, k):
        # a x + b (y + cdt k) over the (q, c) pair
        return tuple(a * xi + b * (yi + cdt * ki) for xi, yi, ki in zip(x, y, k))

    s0 = (q, c)
    k1 = rhs(s0)
    y1 = comb(0.0, s0, 1.0, s0, p.dt, k1)
    k2 = rhs(y1)
    y2 = comb(0.75, s0, 0.25, y1, p.dt, k2)
    k3 = rhs(y2)
    qn, cn = comb(1.0 / 3.0, s0, 2.0 / 3.0, y2, p.dt, k3)
    return qn, cn, last["psi"]
