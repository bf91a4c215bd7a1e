"""Pin the oracle functions that no golden file covers (CPU only).

Each family has one check function built from closed forms, independent
library routines and the reference's own test cases; the oracle must pass it
and a plausible planted mistake in a test-local copy of the oracle function
(a dropped term, a wrong sign or coefficient, a shifted index) must fail it.

* ``arakawa_jacobian``: reference ``tests/test_jacobian.py:94-230``
  (J(x, y) = 1, J(xy, x^2) = -2x^2 within 5 dx^2, antisymmetry, J(f, f) = 0,
  constant arguments, periodic conservation of the integral, energy and
  enstrophy).
* TVD limiters: reference ``tests/test_limiters.py:11-106`` values, the
  symmetry phi(1/r) = phi(r)/r and Sweby's TVD region.
* WENO3/WENO5 and their WENO-Z forms (``advection/weno.py:297-522``):
  constants and linear data exact (reference ``tests/test_weno.py``), the
  left/right mirror identity, boundedness at a step, and the formal order of
  accuracy on cell averages of a smooth function.
* naive / upwind1 / upwind2 / upwind3 / TVD face fluxes
  (``advection/reconstruction.py:582-1102``): linear exactness, order of
  accuracy, mirror symmetry under x -> -x, the boundary fallbacks.
* ``rk3_ssp_step`` (``timestepping/explicit_rk.py:74-121``): the stability
  polynomial 1 + z + z^2/2 + z^3/6 on y' = lambda y, third-order convergence
  and the reference's 1e-6 accuracy at dt = 0.01
  (``tests/test_timestepping.py:60-106``).
* ``enforce_periodic`` (``boundary/boundary.py:52-78``): equal to numpy's
  wrap padding of the interior for every ghost width.
* Advection2D + rk3_ssp_step on a periodic grid: order > 2 for WENO5 under
  negative flow (reference ``tests/test_advection.py:558-628``), constant
  fields give zero tendency for every method, masked and unmasked.
"""

import numpy as np
import pytest

from oracle import fvx

# ---------------------------------------------------------------------------
# Arakawa Jacobian
# ---------------------------------------------------------------------------


def _periodic(ny, nx, func):
    """Field on (ny, nx) with a one-cell periodic ghost ring (interior period)."""
    j = np.arange(ny, dtype=float)[:, None] / (ny - 2)
    i = np.arange(nx, dtype=float)[None, :] / (nx - 2)
    f = func(j + 0 * i, i + 0 * j)
    f[0, :] = f[ny - 2, :]
    f[ny - 1, :] = f[1, :]
    f[:, 0] = f[:, nx - 2]
    f[:, nx - 1] = f[:, 1]
    return f


def _arakawa_failures(jac):
    fails = []
    # J(x, y) = 1 exactly, square and non-square spacing
    for dx, dy in ((1.0, 1.0), (0.5, 1.0)):
        Y, X = np.meshgrid(np.arange(20) * dy, np.arange(20) * dx, indexing="ij")
        if not np.allclose(jac(X, Y, dx, dy), 1.0, rtol=0, atol=1e-12):
            fails.append(f"J(x,y)=1 dx={dx}")
    # J(xy, x^2) = -2 x^2 (second order)
    d = 0.1
    Y, X = np.meshgrid(np.arange(30) * d, np.arange(30) * d, indexing="ij")
    if not np.allclose(jac(X * Y, X**2, d, d), -2.0 * X[1:-1, 1:-1] ** 2, rtol=0, atol=5 * d**2):
        fails.append("J(xy,x^2)")
    # antisymmetry, J(f, f) = 0, constants
    Y, X = np.meshgrid(np.arange(12.0), np.arange(10.0), indexing="ij")
    f, g = np.sin(X) * np.cos(Y), np.cos(X) * np.sin(Y) + 0.3 * X * Y
    if not np.allclose(jac(f, g, 1.0, 1.0), -jac(g, f, 1.0, 1.0), rtol=0, atol=1e-13):
        fails.append("antisymmetry")
    if not np.allclose(jac(X**2 + Y**2, X**2 + Y**2, 0.5, 1.0), 0.0, atol=1e-12):
        fails.append("J(f,f)")
    if not np.allclose(jac(f, np.full_like(f, 3.14), 1.0, 1.0), 0.0, atol=1e-13):
        fails.append("J(f,c)")
    if not np.allclose(jac(np.full_like(f, 2.71), g, 1.0, 1.0), 0.0, atol=1e-13):
        fails.append("J(c,g)")
    # periodic conservation: integral, energy, enstrophy
    for ny, nx, dx, dy in ((12, 10, 1.0, 1.0), (14, 12, 0.5, 1.0)):
        f = _periodic(ny, nx, lambda j, i: np.sin(2 * np.pi * j) * np.cos(2 * np.pi * i)
                      + 0.2 * np.cos(4 * np.pi * i))
        g = _periodic(ny, nx, lambda j, i: np.cos(2 * np.pi * j) + np.sin(2 * np.pi * i)
                      + 0.1 * np.sin(2 * np.pi * (i + j)))
        J = jac(f, g, dx, dy)
        scale = np.abs(J).sum()
        for name, w in (("integral", 1.0), ("energy", f[1:-1, 1:-1]), ("enstrophy", g[1:-1, 1:-1])):
            if abs(np.sum(w * J)) > 1e-12 * scale:
                fails.append(f"{name} conservation {ny}x{nx}")
    return fails


def _arakawa_mutant(kind):
    """Test-local copies of the oracle's Jacobian with one planted mistake."""

    def jac(f, g, dx, dy):
        E, W = (slice(1, -1), slice(2, None)), (slice(1, -1), slice(None, -2))
        N, S = (slice(2, None), slice(1, -1)), (slice(None, -2), slice(1, -1))
        NE, NW = (slice(2, None), slice(2, None)), (slice(2, None), slice(None, -2))
        SE, SW = (slice(None, -2), slice(2, None)), (slice(None, -2), slice(None, -2))
        jpp = (f[E] - f[W]) * (g[N] - g[S]) - (f[N] - f[S]) * (g[E] - g[W])
        s = -1.0 if kind == "sign" else 1.0
        jpx = (f[E] * (g[NE] - g[SE]) - f[W] * (g[NW] - g[SW])
               - f[N] * (g[NE] - g[NW]) + s * f[S] * (g[SE] - g[SW]))
        ne = NW if kind == "index" else NE
        jxp = (g[N] * (f[ne] - f[NW]) - g[S] * (f[SE] - f[SW])
               - g[E] * (f[NE] - f[SE]) + g[W] * (f[NW] - f[SW]))
        if kind == "drop":
            return (jpp + jpx) / (8.0 * dx * dy)
        return (jpp + jpx + jxp) / (12.0 * dx * dy)

    if kind == "axes":  # rows taken as x and columns as y
        base = _arakawa_mutant("none")
        return lambda f, g, dx, dy: base(np.swapaxes(f, -1, -2), np.swapaxes(g, -1, -2), dy,
                                         dx).swapaxes(-1, -2)
    return jac


def test_arakawa_pins():
    assert _arakawa_failures(fvx.arakawa_jacobian) == []


@pytest.mark.parametrize("kind", ["drop", "sign", "index", "axes"])
def test_arakawa_planted_mistake_is_caught(kind):
    assert _arakawa_failures(_arakawa_mutant(kind)), kind


def test_arakawa_batched_matches_loop():
    rng = np.random.default_rng(0)
    f, g = rng.standard_normal((2, 3, 9, 11))
    out = fvx.arakawa_jacobian(f, g, 0.3, 0.7)
    assert out.shape == (3, 7, 9)
    for b in range(3):
        np.testing.assert_array_equal(out[b], fvx.arakawa_jacobian(f[b], g[b], 0.3, 0.7))


# ---------------------------------------------------------------------------
# TVD limiters
# ---------------------------------------------------------------------------

LIMITER_VALUES = {  # reference tests/test_limiters.py:11-106
    "minmod": [(-2.0, 0.0), (0.0, 0.0), (0.3, 0.3), (0.7, 0.7), (1.0, 1.0), (1.5, 1.0), (5.0, 1.0)],
    "van_leer": [(-1.0, 0.0), (0.0, 0.0), (1.0, 1.0), (3.0, 1.5)],
    "superbee": [(-1.0, 0.0), (0.0, 0.0), (0.5, 1.0), (1.0, 1.0), (3.0, 2.0), (10.0, 2.0),
                 (0.25, 0.5), (1.5, 1.5)],
    "mc": [(-1.0, 0.0), (0.0, 0.0), (0.1, 0.2), (1.0, 1.0), (2.0, 1.5), (5.0, 2.0)],
}


def _limiter_failures(name, phi):
    fails = []
    for r, want in LIMITER_VALUES[name]:
        if abs(float(phi(np.array(r))) - want) > 1e-12:
            fails.append(f"{name}({r})")
    r = np.concatenate([np.linspace(1e-3, 10, 997), [100.0, 1000.0]])
    p = phi(r)
    # Sweby's second-order TVD region: r <= phi <= min(2r, 1)... bounded by
    # 0 <= phi <= min(2 r, 2), phi(r) = 0 for r <= 0, and symmetry.
    if np.any(p < -1e-15) or np.any(p > np.minimum(2 * r, 2.0) + 1e-12):
        fails.append(f"{name} TVD region")
    if np.any(phi(-r) != 0.0):
        fails.append(f"{name} negative r")
    if not np.allclose(phi(1.0 / r), p / r, rtol=1e-12, atol=1e-15):
        fails.append(f"{name} symmetry")
    return fails


@pytest.mark.parametrize("name", sorted(fvx.LIMITERS))
def test_limiter_pins(name):
    assert _limiter_failures(name, fvx.LIMITERS[name]) == []


@pytest.mark.parametrize("name,mutant", [
    ("minmod", lambda r: np.minimum(1.0, r)),                               # no max(0, .)
    ("van_leer", lambda r: (r + np.abs(r)) / (1.0 + r * r)),               # wrong denominator
    ("superbee", lambda r: np.maximum(0.0, np.maximum(np.minimum(2 * r, 1.0), np.minimum(r, 1.0)))),
    ("mc", lambda r: np.maximum(0.0, np.minimum(np.minimum(1.0 + r, 2.0 * r), 2.0))),
])
def test_limiter_planted_mistake_is_caught(name, mutant):
    assert _limiter_failures(name, mutant), name


# ---------------------------------------------------------------------------
# WENO point reconstructions
# ---------------------------------------------------------------------------


def _cell_averages(F, x_edges):
    """Exact cell averages of f from its antiderivative F."""
    return (F(x_edges[1:]) - F(x_edges[:-1])) / np.diff(x_edges)


def _order_left(rec, width, n_list=(40, 80)):
    """Observed order of a left-biased reconstruction of exp(x) on [0, 1] from
    cell averages; the face value sits right of cell ``width // 2``."""
    errs = []
    for n in n_list:
        e = np.linspace(0.0, 1.0, n + 1)
        a = _cell_averages(np.exp, e)
        s = [a[k:n - width + 1 + k] for k in range(width)]
        face = e[width // 2 + 1: n - width + width // 2 + 2]
        errs.append(np.max(np.abs(rec(*s) - np.exp(face))))
    return np.log(errs[0] / errs[1]) / np.log(n_list[1] / n_list[0])


def _order_right(rec, width, n_list=(40, 80)):
    """Observed order of a right-biased reconstruction (stencil q_{i-w/2+1}..);
    the face value sits right of cell ``(width - 1) // 2 - 1 + 1``."""
    errs = []
    for n in n_list:
        e = np.linspace(0.0, 1.0, n + 1)
        a = _cell_averages(np.exp, e)
        s = [a[k:n - width + 1 + k] for k in range(width)]
        off = 1 if width == 3 else 2  # face between q0 and qp
        face = e[off: n - width + 1 + off]
        errs.append(np.max(np.abs(rec(*s) - np.exp(face))))
    return np.log(errs[0] / errs[1]) / np.log(n_list[1] / n_list[0])


WENO = {
    # name: (left, right, width, order, mirror of left)
    "weno3": (fvx.weno_3pts, fvx.weno_3pts_right, 3, 3),
    "wenoz3": (fvx.wenoz_3pts, fvx.wenoz_3pts_right, 3, 3),
    "weno5": (fvx.weno_5pts, fvx.weno_5pts_right, 5, 5),
    "wenoz5": (fvx.wenoz_5pts, fvx.wenoz_5pts_right, 5, 5),
}


def _weno_failures(left, right, width, order):
    fails = []
    rng = np.random.default_rng(3)
    c = 2.5
    const = [np.full(7, c)] * width
    if not (np.allclose(left(*const), c, rtol=1e-12) and np.allclose(right(*const), c, rtol=1e-12)):
        fails.append("constant")
    # linear data: every candidate is exact, so the face value is exact
    dx = 0.05
    q = np.arange(24, dtype=float) * dx + 0.3
    s = [q[k:len(q) - width + 1 + k] for k in range(width)]
    m = np.arange(len(s[0]))
    lface = (m + width // 2 + 0.5) * dx + 0.3
    rface = (m + (0 if width == 3 else 1) + 0.5) * dx + 0.3
    if not np.allclose(left(*s), lface, rtol=0, atol=1e-12):
        fails.append("linear left")
    if not np.allclose(right(*s), rface, rtol=0, atol=1e-12):
        fails.append("linear right")
    # mirror identity: right-biased(q) == left-biased(reversed stencil)
    z = [rng.standard_normal(50) for _ in range(width)]
    if not np.allclose(right(*z), left(*z[::-1]), rtol=1e-12, atol=1e-14):
        fails.append("mirror")
    # bounded at a step (ENO property) and upwind selection
    for lo, hi in ((0.0, 1.0), (-1.0, 2.0)):
        st = [np.array(lo if k <= width // 2 else hi) for k in range(width)]
        v = float(left(*st))
        if not (lo - 1e-9 <= v <= hi + 1e-9):
            fails.append(f"bounded {lo},{hi}")
    if float(right(*[np.array(0.0 if k < width // 2 else 1.0) for k in range(width)])) < 0.9:
        fails.append("right upwind at step")
    # formal order on cell averages of a smooth function
    if _order_left(left, width) < order - 0.4:
        fails.append("order left")
    if _order_right(right, width) < order - 0.4:
        fails.append("order right")
    return fails


@pytest.mark.parametrize("name", sorted(WENO))
def test_weno_pins(name):
    assert _weno_failures(*WENO[name]) == []


def _weno5_swapped_weights(qmm, qm, q0, qp, qpp):
    qi1, qi2, qi3, b1, b2, b3 = fvx._weno5_parts(qmm, qm, q0, qp, qpp)
    w1, w2, w3 = 0.3 / (b1 + 1e-8) ** 2, 0.6 / (b2 + 1e-8) ** 2, 0.1 / (b3 + 1e-8) ** 2
    return (w1 * qi1 + w2 * qi2 + w3 * qi3) / (w1 + w2 + w3)


def _weno3_wrong_candidate(qm, q0, qp):
    b1, b2 = (q0 - qm) ** 2, (qp - q0) ** 2
    w1, w2 = (1 / 3) / (b1 + 1e-8) ** 2, (2 / 3) / (b2 + 1e-8) ** 2
    return (w1 * (-0.5 * qm + 1.5 * q0) + w2 * (0.5 * q0 + 0.6 * qp)) / (w1 + w2)


def _wenoz5_wrong_tau(qmm, qm, q0, qp, qpp):
    qi1, qi2, qi3, b1, b2, b3 = fvx._weno5_parts(qmm, qm, q0, qp, qpp)
    tau = np.abs(b1 - b2)  # tau5 is |b1 - b3|
    w1, w2, w3 = (0.1 * (1 + tau / (b + 1e-16)) for b in (b1,)), 0, 0
    w1 = 0.1 * (1 + tau / (b1 + 1e-16))
    w2 = 0.6 * (1 + tau / (b2 + 1e-16))
    w3 = 0.3 * (1 + tau / (b3 + 1e-16))
    return (w1 * qi1 + w2 * qi2 + w3 * qi3) / (w1 + w2 + w3)


@pytest.mark.parametrize("name,mutant_left", [
    ("weno5", _weno5_swapped_weights),
    ("weno3", _weno3_wrong_candidate),
    ("wenoz5", _wenoz5_wrong_tau),
])
def test_weno_planted_mistake_is_caught(name, mutant_left):
    _, right, width, order = WENO[name]
    assert _weno_failures(mutant_left, right, width, order), name


def test_weno_smoothness_indicators_vanish_on_constants():
    c = np.full(4, 1.7)
    *_, b1, b2, b3 = fvx._weno5_parts(c, c, c, c, c)
    assert np.allclose([b1, b2, b3], 0.0, rtol=0, atol=1e-28)


# ---------------------------------------------------------------------------
# Face fluxes of Advection2D (reconstruction on the x axis; y is its transpose)
# ---------------------------------------------------------------------------

FLUX_ORDER = {"naive": 2, "upwind1": 1, "upwind2": 2, "upwind3": 3, "weno3": 3, "wenoz3": 3,
              "weno5": 5, "wenoz5": 5, "tvd:minmod": 2, "tvd:van_leer": 2, "tvd:superbee": 2,
              "tvd:mc": 2}


def _flux(method, h, u):
    if method.startswith("tvd:"):
        return fvx.face_flux_x(h, u, "tvd", method[4:])
    return fvx.face_flux_x(h, u, method)


def _face_order(method, sign, n_list=(64, 128)):
    """Observed order of the east-face value (flux / u) reconstructed from
    cell averages of exp(x), over faces whose full stencil lies inside."""
    errs = []
    for n in n_list:
        e = np.linspace(0.0, 1.0, n + 1)
        a = _cell_averages(np.exp, e)
        h = np.tile(a, (5, 1))
        u = np.full_like(h, float(sign))
        f = _flux(method, h, u)[2, 4:-4] / sign
        errs.append(np.max(np.abs(f - np.exp(e[5:-4]))))
    return np.log(errs[0] / errs[1]) / np.log(n_list[1] / n_list[0])


@pytest.mark.parametrize("method", sorted(FLUX_ORDER))
@pytest.mark.parametrize("sign", [1, -1])
def test_face_flux_order(method, sign):
    order = FLUX_ORDER[method]
    # limited TVD schemes clip at the smooth extremum-free data only mildly
    assert _face_order(method, sign) > order - (0.5 if order > 1 else 0.2), (method, sign)


@pytest.mark.parametrize("method", sorted(FLUX_ORDER))
def test_face_flux_linear_exact_and_mirror(method):
    """Linear data: the face value is exact wherever the full stencil fits
    (every scheme above first order reproduces linear cell averages); and
    reversing x with u -> -u mirrors the flux in the deep interior."""
    n = 20
    x = np.arange(n, dtype=float)
    h = np.tile(0.7 * x + 2.0, (4, 1))
    for sign in (1, -1) if method != "upwind1" else ():
        u = np.full_like(h, 0.5 * sign)
        f = _flux(method, h, u)[1, 3:-3] / (0.5 * sign)
        exact = 0.7 * (x[3:-3] + 0.5) + 2.0
        assert np.allclose(f, exact, rtol=0, atol=1e-6 if method.startswith("tvd") else 1e-12), \
            (method, sign)
    rng = np.random.default_rng(5)
    h = rng.standard_normal((6, n))
    u = rng.standard_normal((6, n))
    f = _flux(method, h, u)
    hr = h[:, ::-1].copy()
    ur = np.zeros_like(u)
    ur[:, :-1] = -u[:, ::-1][:, 1:]  # face i+1/2 -> face (n-2-i)+1/2
    fr = _flux(method, hr, ur)
    # face i (east of cell i) <-> face n-2-i of the reversed grid
    i = np.arange(4, n - 5)
    assert np.allclose(fr[1:-1, n - 2 - i], -f[1:-1, i], rtol=1e-12, atol=1e-12), method


def test_face_flux_boundary_fallbacks():
    """Faces next to the walls fall back as reconstruction.py:832-864 does:
    positive flow WENO3 at faces 1 and Nx-2; negative flow WENO3-right at
    Nx-3 and upwind1 (the value east of the face) at Nx-2."""
    rng = np.random.default_rng(6)
    h = rng.standard_normal((5, 12))
    N = h.shape[1]
    fp = fvx.face_flux_x(h, np.ones_like(h), "weno5")[2]
    assert np.isclose(fp[1], fvx.weno_3pts(h[2, 0], h[2, 1], h[2, 2]))
    assert np.isclose(fp[N - 2], fvx.weno_3pts(h[2, N - 3], h[2, N - 2], h[2, N - 1]))
    assert np.isclose(fp[5], fvx.weno_5pts(*h[2, 3:8]))
    fn = fvx.face_flux_x(h, -np.ones_like(h), "weno5")[2]
    assert np.isclose(fn[N - 2], -h[2, N - 1])
    assert np.isclose(fn[N - 3], -fvx.weno_3pts_right(h[2, N - 3], h[2, N - 2], h[2, N - 1]))
    assert np.isclose(fn[1], -fvx.weno_5pts_right(*h[2, 0:5]))
    assert np.all(fp[0] == 0) and np.all(fp[N - 1] == 0)
    ft = fvx.face_flux_x(h, -np.ones_like(h), "tvd", "minmod")[2]
    assert np.isclose(ft[N - 2], -h[2, N - 1])


def _upwind3_mutant(h, u):
    """upwind3 with the outer coefficients swapped."""
    H = h[1:-1]
    pos = 1.0 / 3.0 * H[:, :-2] + 5.0 / 6.0 * H[:, 1:-1] - 1.0 / 6.0 * H[:, 2:]
    neg = np.concatenate([1.0 / 3.0 * H[:, 1:-2] + 5.0 / 6.0 * H[:, 2:-1] - 1.0 / 6.0 * H[:, 3:],
                          H[:, -1:]], axis=1)
    return fvx.interior(np.where(u[1:-1, 1:-1] >= 0, pos, neg) * u[1:-1, 1:-1], h.shape)


def test_face_flux_planted_mistake_is_caught():
    n = 20
    x = np.arange(n, dtype=float)
    h = np.tile(0.7 * x + 2.0, (4, 1))
    u = np.ones_like(h)
    f = _upwind3_mutant(h, u)[1, 3:-3]
    assert not np.allclose(f, 0.7 * (x[3:-3] + 0.5) + 2.0, atol=1e-12)


# ---------------------------------------------------------------------------
# rk3_ssp_step
# ---------------------------------------------------------------------------


def _rk3_failures(step):
    fails = []
    # stability polynomial on y' = lam y: R(z) = 1 + z + z^2/2 + z^3/6
    for lam, dt in ((-1.0, 0.1), (-3.0, 0.5), (2.0, 0.01), (-1.0, 1.7)):
        z = lam * dt
        got = step(np.array([1.0]), lambda y: lam * y, dt)[0]
        if abs(got - (1 + z + z * z / 2 + z**3 / 6)) > 1e-14 * max(1.0, abs(got)):
            fails.append(f"R(z) lam={lam} dt={dt}")
    # third-order convergence on a nonlinear ODE y' = -y^2 (y = 1 / (1 + t))
    errs = []
    for n in (10, 20, 40):
        y = np.array([1.0])
        for _ in range(n):
            y = step(y, lambda v: -v * v, 1.0 / n)
        errs.append(abs(y[0] - 0.5))
    rate = np.log2(errs[-2] / errs[-1])
    if rate < 2.85:
        fails.append(f"order {rate:.2f}")
    # reference accuracy: dt = 0.01, 100 steps of y' = -y, rtol 1e-6
    y = np.array([1.0])
    for _ in range(100):
        y = step(y, lambda v: -v, 0.01)
    if abs(y[0] - np.exp(-1.0)) > 1e-6 * np.exp(-1.0):
        fails.append("accuracy")
    return fails


def _rk3_mutant(state, rhs_fn, dt):
    k1 = rhs_fn(state)
    y1 = state + dt * k1
    k2 = rhs_fn(y1)
    y2 = 0.75 * state + 0.25 * y1 + 0.25 * dt * k2
    k3 = rhs_fn(y2)
    return (1.0 / 3.0) * state + (2.0 / 3.0) * y2 + (1.0 / 3.0) * dt * k3  # 2/3 -> 1/3


def test_rk3_pins():
    assert _rk3_failures(fvx.rk3_ssp_step) == []


def test_rk3_planted_mistake_is_caught():
    assert _rk3_failures(_rk3_mutant)


# ---------------------------------------------------------------------------
# enforce_periodic
# ---------------------------------------------------------------------------


def _periodic_failures(fill):
    fails = []
    rng = np.random.default_rng(7)
    for ng in (1, 2, 3, 4):
        interior = rng.standard_normal((9, 13))
        full = np.pad(interior, ng, constant_values=np.nan)
        if not np.array_equal(fill(full, ng), np.pad(interior, ng, mode="wrap")):
            fails.append(f"ng={ng}")
    return fails


def test_enforce_periodic_pins():
    assert _periodic_failures(fvx.enforce_periodic) == []


def test_enforce_periodic_planted_mistake_is_caught():
    def shifted(field, ng=1):
        f = np.array(field, np.float64)
        f[:ng, :] = f[-2 * ng - 1:-ng - 1, :]  # off by one row
        f[-ng:, :] = f[ng:2 * ng, :]
        f[:, :ng] = f[:, -2 * ng:-ng]
        f[:, -ng:] = f[:, ng:2 * ng]
        return f

    assert _periodic_failures(shifted)


# ---------------------------------------------------------------------------
# Advection2D (+ rk3_ssp_step) end to end
# ---------------------------------------------------------------------------

METHODS = ("naive", "upwind1", "upwind2", "upwind3", "weno3", "weno5", "wenoz5",
           "minmod", "van_leer", "superbee", "mc")


@pytest.mark.parametrize("method", METHODS)
def test_constant_field_zero_tendency(method):
    """Reference tests/test_advection.py:116-171, 532-555 (both flow signs)."""
    n = 18
    h = np.full((n, n), 1.3)
    for s in (1.0, -1.0):
        u = np.full((n, n), 0.4 * s)
        v = np.full((n, n), -0.7 * s)
        out = fvx.advection_2d(h, u, v, 0.1, 0.2, method)
        assert np.allclose(out[4:-4, 4:-4], 0.0, atol=1e-12), method
        assert np.all(out[:2] == 0) and np.all(out[:, -2:] == 0)


@pytest.mark.parametrize("method", ["weno3", "weno5", "wenoz5", "minmod", "van_leer"])
def test_masked_all_ocean_matches_unmasked_interior(method):
    """Reference tests/test_advection.py:341-354: with every interior cell wet
    the tiered fluxes equal the unmasked ones away from the dry ghost ring."""
    n = 22
    rng = np.random.default_rng(8)
    h, u, v = rng.standard_normal((3, n, n))
    wet = np.zeros((n, n), bool)
    wet[1:-1, 1:-1] = True
    m = fvx.Mask2D(wet)
    a = fvx.advection_2d(h, u, v, 0.1, 0.1, method, m)
    b = fvx.advection_2d(h, u, v, 0.1, 0.1, method)
    np.testing.assert_allclose(a[5:-5, 5:-5], b[5:-5, 5:-5], rtol=1e-13, atol=1e-13)


def _periodic_advect_error(n, method, sign=-1.0, ng=4, periods=1):
    """sin(2 pi x) advected once around a periodic domain by rk3_ssp_step
    (reference tests/test_advection.py:558-628, in 2-D with v = 0)."""
    dx = 1.0 / n
    N = n + 2 * ng
    Ny = 2 * ng + 2  # the field is constant along y: a periodic strip
    x = (np.arange(N) - ng + 0.5) * dx
    q0 = np.tile(np.sin(2 * np.pi * x), (Ny, 1))
    q0 = fvx.enforce_periodic(q0, ng)
    u = np.full((Ny, N), sign)
    v = np.zeros((Ny, N))
    dt = 0.3 * dx
    nsteps = int(round(periods / dt))
    dt = periods / nsteps
    q = q0.copy()
    for _ in range(nsteps):
        q = fvx.rk3_ssp_step(q, lambda s: fvx.advection_2d(fvx.enforce_periodic(s, ng), u, v,
                                                            dx, dx, method), dt)
        q = fvx.enforce_periodic(q, ng)
    return np.max(np.abs(q[ng:-ng, ng:-ng] - q0[ng:-ng, ng:-ng]))


@pytest.mark.parametrize("method,min_order", [("upwind1", 0.8), ("weno3", 1.2), ("weno5", 2.0)])
def test_periodic_advection_convergence(method, min_order):
    e1 = _periodic_advect_error(32, method)
    e2 = _periodic_advect_error(128, method)
    assert np.log(e1 / e2) / np.log(4.0) > min_order, (method, e1, e2)


# ---------------------------------------------------------------------------
# Vertical modes (reference tests/test_vertical_modes.py)
# ---------------------------------------------------------------------------


def _coupling_failures(build):
    fails = []
    A1 = build([1000.0], [9.81])
    if A1.shape != (1, 1) or abs(A1[0, 0] - 1.0 / (1000.0 * 9.81)) > 1e-18:
        fails.append("single layer")
    H0 = H1 = 500.0
    g0 = g1 = 0.02
    A2 = build([H0, H1], [g0, g1])
    want2 = [[1 / (H0 * g0) + 1 / (H0 * g1), -1 / (H0 * g1)], [-1 / (H1 * g1), 1 / (H1 * g1)]]
    if not np.allclose(A2, want2, rtol=1e-14, atol=0):
        fails.append("two layers")
    g = [0.02, 0.015, 0.015]
    A3 = build([300.0, 400.0, 300.0], g)
    row1 = [-1 / (400.0 * g[1]), (1 / 400.0) * (1 / g[2] + 1 / g[1]), -1 / (400.0 * g[2])]
    if not np.allclose(A3[1], row1, rtol=1e-14, atol=0):
        fails.append("three-layer interior row")
    if A3[0, 2] != 0 or A3[2, 0] != 0 or not np.all(np.diag(A3) > 0):
        fails.append("tridiagonal, positive diagonal")
    # every row of A sums to zero except the top one (1/(H0 g0)): the
    # barotropic mode has eigenvalue ~ 1/(H g0) only through the top row
    if not np.allclose(A3[1:].sum(axis=1), 0.0, atol=1e-15):
        fails.append("row sums")
    return fails


def test_vertical_mode_pins():
    assert _coupling_failures(fvx.build_coupling_matrix) == []
    f0 = 1e-4
    # single layer: Rd = sqrt(H g') / |f0|
    r, _, _ = fvx.decompose_vertical_modes(fvx.build_coupling_matrix([1000.0], [9.81]), f0)
    assert r[0] == pytest.approx(np.sqrt(1000.0 * 9.81) / f0, rel=1e-14)
    # two equal layers: A = (1 / (H g)) [[2, -1], [-1, 1]], eigenvalues (3 -+ sqrt 5) / 2
    A = fvx.build_coupling_matrix([500.0, 500.0], [0.02, 0.02])
    r, l2m, m2l = fvx.decompose_vertical_modes(A, f0)
    lam = np.array([(3 - np.sqrt(5)) / 2, (3 + np.sqrt(5)) / 2]) / (500.0 * 0.02)
    np.testing.assert_allclose(np.sort(r), np.sort(1.0 / (f0 * np.sqrt(lam))), rtol=1e-13)
    np.testing.assert_allclose(l2m @ m2l, np.eye(2), atol=1e-14)
    np.testing.assert_allclose(m2l @ np.diag(1.0 / (f0 * r) ** 2) @ l2m, A, rtol=1e-12)
    field = np.random.default_rng(9).standard_normal((2, 5, 6))
    np.testing.assert_allclose(fvx.mode_to_layer(fvx.layer_to_mode(field, l2m), m2l), field,
                               atol=1e-13)


def test_vertical_mode_planted_mistake_is_caught():
    def wrong(H, g_prime):
        A = fvx.build_coupling_matrix(H, g_prime)
        if len(H) > 2:
            A[1, 1] = (1.0 / H[1]) * (1.0 / g_prime[1] + 1.0 / g_prime[0])  # g[i+1] -> g[i-1]
        return A

    assert _coupling_failures(wrong)


# ---------------------------------------------------------------------------
# euler_step / heun_step / rk4_step (timestepping/explicit_rk.py:22-168)
# ---------------------------------------------------------------------------

RK_POLY = {  # stability polynomial R(z) of each scheme on y' = lambda y
    "euler_step": lambda z: 1 + z,
    "heun_step": lambda z: 1 + z + z * z / 2,
    "rk4_step": lambda z: 1 + z + z * z / 2 + z**3 / 6 + z**4 / 24,
}
RK_ORDER = {"euler_step": 1, "heun_step": 2, "rk4_step": 4}


def _rk_failures(name, step):
    fails = []
    for lam, dt in ((-1.0, 0.1), (-3.0, 0.5), (2.0, 0.01)):
        got = step(np.array([1.0]), lambda y: lam * y, dt)[0]
        if abs(got - RK_POLY[name](lam * dt)) > 1e-14 * max(1.0, abs(got)):
            fails.append(f"R(z) lam={lam}")
    errs = []
    for n in (8, 16, 32):
        y = np.array([1.0])
        for _ in range(n):
            y = step(y, lambda v: -v * v, 1.0 / n)
        errs.append(abs(y[0] - 0.5))
    if np.log2(errs[-2] / errs[-1]) < RK_ORDER[name] - 0.15:
        fails.append("order")
    return fails


@pytest.mark.parametrize("name", sorted(RK_POLY))
def test_explicit_rk_pins(name):
    assert _rk_failures(name, getattr(fvx, name)) == []


def test_heun_planted_mistake_is_caught():
    def wrong(state, rhs_fn, dt):  # corrector evaluated at the predictor's start
        k1 = rhs_fn(state)
        return state + 0.5 * dt * (k1 + rhs_fn(state + 0.5 * dt * k1))

    assert _rk_failures("heun_step", wrong)


# ---------------------------------------------------------------------------
# MomentumAdvection2D (diffusion/momentum.py; energy scheme also pinned by the
# reference goldens in test_oracle_golden.py)
# ---------------------------------------------------------------------------


def _momentum_failures(madv):
    fails = []
    n = 14
    dx, dy = 0.3, 0.4
    for scheme in ("energy", "enstrophy", "al"):
        # uniform flow and zero flow: zero tendency (tests/test_momentum.py:50-63, 142-150)
        du, dv = madv(np.full((n, n), 0.7), np.full((n, n), -0.2), dx, dy, scheme)
        if not (np.allclose(du, 0, atol=1e-13) and np.allclose(dv, 0, atol=1e-13)):
            fails.append(f"uniform {scheme}")
    rng = np.random.default_rng(17)
    u, v = rng.standard_normal((2, n, n))
    e = madv(u, v, dx, dy, "energy")
    z = madv(u, v, dx, dy, "enstrophy")
    a = madv(u, v, dx, dy, "al")
    # Arakawa-Lamb = 1/3 energy + 2/3 enstrophy (tests/test_momentum.py:93-106)
    for k in (0, 1):
        if not np.allclose(a[k], e[k] / 3 + 2 * z[k] / 3, rtol=1e-12, atol=1e-12):
            fails.append("al blend")
    # solid-body rotation u = -y, v = x (tests/test_momentum.py:108-140):
    # zeta = 2 everywhere, and with K = (u_T^2 + v_T^2)/2 the energy-scheme
    # tendencies reduce to du = 2 x - x = x at U points and dv = y at V points
    y = np.arange(n, dtype=float) * dy
    x = np.arange(n, dtype=float) * dx
    ur = np.broadcast_to(-y[:, None], (n, n)).copy()
    vr = np.broadcast_to(x[None, :], (n, n)).copy()
    du, dv = madv(ur, vr, dx, dy, "energy")
    if not np.allclose(du[2:-2, 2:-2], np.broadcast_to(x[None, 2:-2] + 0.5 * dx, (n - 4, n - 4)),
                       rtol=1e-12, atol=1e-12):
        fails.append("solid-body du")
    if not np.allclose(dv[2:-2, 2:-2], np.broadcast_to(y[2:-2, None] + 0.5 * dy, (n - 4, n - 4)),
                       rtol=1e-12, atol=1e-12):
        fails.append("solid-body dv")
    return fails


def test_momentum_pins():
    assert _momentum_failures(fvx.momentum_advection_2d) == []


def test_momentum_planted_mistake_is_caught():
    def wrong_k(u, v, dx, dy, scheme="energy", mask=None):  # K without the 1/2
        du, dv = fvx.momentum_advection_2d(u, v, dx, dy, scheme, mask)
        uT, vT = fvx.U_to_T(u), fvx.V_to_T(v)
        K = fvx.interior(0.5 * (uT[1:-1, 1:-1] ** 2 + vT[1:-1, 1:-1] ** 2), u.shape)
        extra = fvx.diff_x_T_to_U(K, dx)
        return du - fvx.interior(extra[2:-2, 2:-2], u.shape, ghost=2), dv

    assert _momentum_failures(wrong_k)


# ---------------------------------------------------------------------------
# Multi-layer QG step (composition of vertical/vertical_modes.py:50-263 with
# the 1.5-layer step and Advection2D)
# ---------------------------------------------------------------------------


def _basin(ny, nx):
    wet = np.ones((ny, nx))
    wet[3:5, 4:7] = 0.0  # an island
    wet[0, :3] = 0.0     # a notch in the coast
    return wet


def _multilayer_failures(invert, rhs, velocity):
    fails = []
    f0 = 9.375e-5
    ny, nx = 10, 12
    L = 1.2e6
    wet = _basin(ny, nx)
    # equal thicknesses: A symmetric positive definite, so the modal solves
    # must reproduce the textbook layer equations lap psi_l - f0^2 (A psi)_l = q_l
    p = fvx.MultiLayerParams(nx, ny, L, L * ny / nx, [500.0, 500.0, 500.0], [9.81, 0.02, 0.02],
                             f0=f0)
    rng = np.random.default_rng(23)
    q = np.zeros((3, ny + 2, nx + 2))
    q[:, 1:-1, 1:-1] = rng.standard_normal((3, ny, nx)) * 1e-6 * wet
    psi = invert(q, p, wet)
    A = fvx.build_coupling_matrix(p.H, p.g_prime)
    lap, idx = fvx.dirichlet_helmholtz_matrix(wet, p.dx, p.dy, 0.0)
    m = wet > 0
    for l in range(3):
        lhs = lap @ psi[l, 1:-1, 1:-1][m] - f0**2 * sum(A[l, k] * psi[k, 1:-1, 1:-1][m]
                                                         for k in range(3))
        if not np.allclose(lhs, q[l, 1:-1, 1:-1][m], rtol=1e-9, atol=1e-9 * np.abs(q).max()):
            fails.append(f"layer equation {l}")
    if np.any(psi[:, 1:-1, 1:-1][:, ~m] != 0) or np.any(psi[:, 0] != 0):
        fails.append("psi on land")
    # forcing enters the top layer only
    F = p.forcing()
    r = rhs(np.zeros_like(q), np.zeros_like(q), p, F, wet)
    if not (np.allclose(r[0, 1:-1, 1:-1], F[1:-1, 1:-1] * wet, rtol=1e-14, atol=0)
            and np.all(r[1:] == 0)):
        fails.append("forcing layer")
    # tracer velocity: uniform for a linear streamfunction psi = -U0 y,
    # discretely non-divergent for any streamfunction
    U0 = 0.7
    yc = p.y_centres()
    lin = np.broadcast_to((-U0 * yc)[:, None], (ny + 2, nx + 2)).copy()
    u, v = velocity(lin, p)
    if not (np.allclose(u[2:-1, 1:-1], U0, rtol=1e-12) and np.allclose(v[1:-1, 2:-1], 0.0,
                                                                        atol=1e-15)):
        fails.append("uniform flow")
    u, v = velocity(rng.standard_normal((ny + 2, nx + 2)), p)
    div = (u[2:-2, 2:-2] - u[2:-2, 1:-3]) / p.dx + (v[2:-2, 2:-2] - v[1:-3, 2:-2]) / p.dy
    if not np.allclose(div, 0.0, atol=1e-12 / p.dx):
        fails.append("divergence")
    return fails


def test_multilayer_pins():
    assert _multilayer_failures(fvx.multilayer_invert, fvx.multilayer_rhs,
                                fvx.tracer_velocity) == []


def test_multilayer_reduces_to_one_and_a_half_layers():
    ny = nx = 12
    L = 1.2e6
    H, g = 800.0, 0.02
    p = fvx.MultiLayerParams(nx, ny, L, L, [H], [g], nu4=1e9, F0=2e-12, dt=1200.0)
    p1 = fvx.QGParams(nx, ny, L, L, Ld=np.sqrt(H * g) / p.f0, nu4=1e9, F0=2e-12, dt=1200.0)
    wet = _basin(ny, nx)
    q = np.zeros((ny + 2, nx + 2))
    q[1:-1, 1:-1] = np.random.default_rng(5).standard_normal((ny, nx)) * 1e-6 * wet
    F = p.forcing()
    got, c, _ = fvx.multilayer_step(q[None], np.zeros_like(q), p, F, wet, fvx.Mask2D(np.pad(wet, 1)))
    want = fvx.qg_step(q, p1, F, wet)
    np.testing.assert_allclose(got[0], want, rtol=1e-10, atol=1e-10 * np.abs(want).max())
    assert np.all(c == 0)


def test_multilayer_planted_mistake_is_caught():
    def swapped(q, p, wet):  # the two mode transforms exchanged
        class P:
            pass
        w = P()
        w.__dict__.update(p.__dict__)
        w.Cl2m, w.Cm2l = p.Cm2l, p.Cl2m
        w.nl = p.nl
        return fvx.multilayer_invert(q, w, wet)

    def bottom_forced(q, psi, p, F, wet):
        return fvx.multilayer_rhs(q, psi, p, F, wet)[::-1].copy()

    def flipped_v(psi, p, mask=None):
        u, v = fvx.tracer_velocity(psi, p, mask)
        return u, -v

    assert "layer equation 0" in _multilayer_failures(swapped, fvx.multilayer_rhs,
                                                      fvx.tracer_velocity)
    assert _multilayer_failures(fvx.multilayer_invert, bottom_forced, fvx.tracer_velocity)
    assert _multilayer_failures(fvx.multilayer_invert, fvx.multilayer_rhs, flipped_v)


# ---------------------------------------------------------------------------
# Prolongation of the Dirichlet multigrid hierarchy (no reference counterpart)
# ---------------------------------------------------------------------------


def _prolong_dirichlet_failures(prolong):
    fails = []
    nc = 8
    mc = np.ones((nc, nc))
    mf = np.ones((2 * nc, 2 * nc))
    # coarse centres X = (I + 1/2) 2h, fine centres x = (i + 1/2) h (h = 1):
    # the 9/3/3/1 weights are bilinear, so a linear field is reproduced on
    # every fine cell whose four coarse neighbours exist
    X = (np.arange(nc) + 0.5) * 2.0
    x = np.arange(2 * nc) + 0.5
    g = 0.3 * X[None, :] - 1.7 * X[:, None] + 2.0
    f = prolong(g, mc, mf)
    want = 0.3 * x[None, :] - 1.7 * x[:, None] + 2.0
    if not np.allclose(f[1:-1, 1:-1], want[1:-1, 1:-1], rtol=0, atol=1e-12):
        fails.append("bilinear interior")
    # at a wall the coarse value beyond it is zero (psi = 0): the first fine
    # cell gets 3/4 of its coarse neighbour's value in 1-D
    c = np.ones((nc, nc))
    e = prolong(c, mc, mf)
    if not (np.isclose(e[5, 0], 0.75) and np.isclose(e[0, 0], 0.75 * 0.75)
            and np.isclose(e[5, 5], 1.0)):
        fails.append("walls")
    # land: a dry coarse cell contributes zero, the fine mask is applied
    mc2 = mc.copy()
    mc2[3, 3] = 0.0
    mf2 = mf.copy()
    mf2[6, 6] = 0.0
    e2 = prolong(c, mc2, mf2)
    if not (e2[6, 6] == 0.0 and np.isclose(e2[6, 7], 7 / 16) and np.isclose(e2[5, 5], 15 / 16)):
        fails.append("land")
    return fails


def test_prolongate_dirichlet_pins():
    assert _prolong_dirichlet_failures(fvx.prolongate_dirichlet) == []


def test_prolongate_dirichlet_planted_mistake_is_caught():
    # the reference's mask-renormalised weights extrapolate to the wall
    assert "walls" in _prolong_dirichlet_failures(fvx.prolongate)


def _restrict_dirichlet_failures(restrict):
    fails = []
    rng = np.random.default_rng(5)
    nc = 8
    mc = np.ones((nc, nc))
    mf = np.ones((2 * nc, 2 * nc))
    # wet everywhere: the mean of the 2x2 children, so a constant is kept and
    # 4 x the coarse sum equals the fine sum (the residual's integral)
    if not np.allclose(restrict(np.full((2 * nc, 2 * nc), 3.0), mf, mc), 3.0, rtol=0, atol=1e-14):
        fails.append("constant")
    v = rng.standard_normal((2 * nc, 2 * nc))
    if not np.isclose(4.0 * restrict(v, mf, mc).sum(), v.sum(), rtol=1e-12):
        fails.append("integral")
    if not np.isclose(restrict(v, mf, mc)[2, 5], v[4:6, 10:12].mean()):
        fails.append("children")
    # land: a dry child counts as zero (psi on land), not renormalised away;
    # a dry coarse cell receives zero
    mf2 = mf.copy()
    mf2[6, 6] = 0.0
    mc2 = mc.copy()
    mc2[0, 7] = 0.0
    r = restrict(np.ones((2 * nc, 2 * nc)), mf2, mc2)
    if not np.isclose(r[3, 3], 0.75):
        fails.append("land")
    if r[0, 7] != 0.0:
        fails.append("coarse mask")
    return fails


def test_restrict_dirichlet_pins():
    assert _restrict_dirichlet_failures(fvx.restrict_dirichlet) == []


def test_restrict_dirichlet_planted_mistake_is_caught():
    # the reference's wet-count renormalisation keeps 1 where a child is dry
    assert _restrict_dirichlet_failures(fvx.restrict) == ["land"]


# ---------------------------------------------------------------------------
# Coarse-level wall correction of the Dirichlet hierarchy (no reference
# counterpart)
# ---------------------------------------------------------------------------


def _wall_failures(g_of_level):
    fails = []
    if g_of_level(0) != 0.0:
        fails.append("level 0")
    # The fine operator's zero ghost puts psi = 0 at x = -h/2 (the ghost
    # centre); a profile linear in x that vanishes there has a zero second
    # difference on every coarse level once the wall-side ghost is g u_0.
    h, n = 1.0 / 64, 16
    for level in (1, 2, 3):
        H = h * 2**level
        x = (np.arange(n) + 0.5) * H
        u = np.tile(x + 0.5 * h, (n, 1))
        lev = fvx.Level(np.ones((n, n)), None, None, H, H, 0.0, None, "dirichlet",
                        g_of_level(level))
        lap = fvx.apply_operator(u, lev)
        if not np.allclose(lap[1:-1, :-1], 0.0, atol=1e-9 * np.abs(u).max() / H**2):
            fails.append(f"linear profile level {level}")
    return fails


def test_wall_correction_pins():
    assert _wall_failures(fvx.wall_g) == []
    # the hierarchy wires it on coarse Dirichlet levels only, diagonal included
    m = np.ones((32, 32))
    levs = fvx.build_multigrid(m, 1.0, 1.0, 0.5, n_levels=3, kind="dirichlet", wall=True)
    assert [lv.wall_g for lv in levs] == [fvx.wall_g(l) for l in range(3)]
    for lv in levs:
        A = fvx.level_matrix(lv)
        assert np.allclose(A.diagonal().reshape(lv.mask.shape) * lv.inv_diag, 1.0)
        e = np.zeros(lv.mask.shape)
        e[0, 3] = 1.0
        assert np.allclose(fvx.apply_operator(e, lv).ravel(), A @ e.ravel())
    plain = fvx.build_multigrid(m, 1.0, 1.0, 0.5, n_levels=3, kind="dirichlet")
    assert all(lv.wall_g == 0.0 for lv in plain)


def test_wall_correction_planted_mistakes_are_caught():
    assert _wall_failures(lambda l: -fvx.wall_g(l))
    assert _wall_failures(lambda l: (1.0 - 2.0**l) / 2.0**l)
    assert _wall_failures(lambda l: 0.0)


# ---------------------------------------------------------------------------
# Chebyshev smoothing weights (no reference counterpart)
# ---------------------------------------------------------------------------


def _chebyshev_failures(weights):
    """The error of m damped-Jacobi sweeps with weights w_k is
    prod_k (1 - w_k D^-1 A) e: for the Chebyshev weights on [lo, 2] that
    polynomial is T_m((2 + lo - 2x) / (2 - lo)) / T_m((2 + lo) / (2 - lo))
    (numpy's Chebyshev basis) at every x of the interval."""
    fails = []
    for lo in (0.05, 0.1, 0.3):
        for m in (1, 2, 3, 5):
            w = np.asarray(weights(lo, m))
            x = np.linspace(lo, 2.0, 41)
            p = np.prod(1.0 - np.outer(x, w), axis=1) if len(w) else np.ones_like(x)
            T = np.polynomial.chebyshev.Chebyshev.basis(m)
            ref = T((2.0 + lo - 2.0 * x) / (2.0 - lo)) / T((2.0 + lo) / (2.0 - lo))
            if len(w) != m or not np.allclose(p, ref, rtol=0.0, atol=1e-12):
                fails.append((lo, m))
    return fails


def _sweep_sequence_failures(jacobi):
    """weighted_jacobi with a weight sequence multiplies the error by
    prod_k (I - w_k D^-1 A) (the level's matrix, land cells held at zero)."""
    rng = np.random.default_rng(5)
    m = (rng.random((7, 6)) > 0.2).astype(np.float64)
    lev = fvx.build_multigrid(m, 0.7, 0.9, 1.5, n_levels=1, kind="dirichlet")[0]
    A = fvx.level_matrix(lev).toarray()
    wet = m.ravel() > 0
    Aw = A[np.ix_(wet, wet)]
    J = np.diag(1.0 / np.diag(Aw)) @ Aw
    x_true = rng.standard_normal(m.shape) * m
    b = fvx.apply_operator(x_true, lev)
    u0 = rng.standard_normal(m.shape) * m
    w = fvx.chebyshev_weights(0.1, 3)
    E = np.eye(J.shape[0])
    for wk in w:
        E = (np.eye(J.shape[0]) - wk * J) @ E
    e = (E @ (u0 - x_true).ravel()[wet])
    got = (jacobi(u0, b, lev, w, 3) - x_true).ravel()[wet]
    return [] if np.allclose(got, e, atol=1e-12 * np.abs(u0).max()) else ["sequence"]


def test_chebyshev_weight_pins():
    assert _chebyshev_failures(fvx.chebyshev_weights) == []
    assert fvx.chebyshev_weights(0.1, 1) == pytest.approx([2.0 / 2.1])
    assert _sweep_sequence_failures(fvx.weighted_jacobi) == []
    # a V-cycle with lo given takes those weights; without, the uniform omega
    m = np.ones((32, 32))
    levs = fvx.build_multigrid(m, 1.0, 1.0, 0.5, n_levels=2, kind="dirichlet")
    b = np.random.default_rng(6).standard_normal((32, 32))
    kw = dict(n_pre=2, n_post=2, n_coarse=3, omega=0.85)
    assert not np.allclose(fvx.v_cycle(levs, 0 * b, b, cheb=0.1, **kw),
                           fvx.v_cycle(levs, 0 * b, b, **kw))


def test_chebyshev_planted_mistakes_are_caught():
    cw = fvx.chebyshev_weights
    assert _chebyshev_failures(lambda lo, m: cw(lo, m, hi=2.2))
    assert _chebyshev_failures(lambda lo, m: cw(0.0, m))
    assert _chebyshev_failures(
        lambda lo, m: list(1.0 / (0.5 * (2 + lo) + 0.5 * (2 - lo) * np.cos(np.pi * np.arange(m) / m))))
    assert _chebyshev_failures(lambda lo, m: cw(lo, m)[:-1])

    def first_weight_only(u, rhs, lev, omega, n):
        return fvx.weighted_jacobi(u, rhs, lev, omega[0], n)

    assert _sweep_sequence_failures(first_weight_only)
