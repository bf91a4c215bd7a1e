"""GPU parity: stencil family, advection, diffusion and the Arakawa Jacobian.

FP64 results must match the oracle to 1e-12 relative (max-norm, scaled by
max|oracle|) -- the reference goldens are checked directly at their own
rtol 1e-12 / atol 1e-14; FP32 results to 1e-5 relative.
"""

import numpy as np
import pytest
import torch

import finitevolx_amd as fx
from oracle import fvx
from tests.conftest import canonical_inputs, load_golden
from tests.gpu_util import assert_close_scaled, need_gpu, random_mask, to_np

pytestmark = pytest.mark.gpu

DIFF_OPS = [
    ("diff_x_T_to_U", "h", fvx.diff_x_T_to_U, "dx"), ("diff_y_T_to_V", "h", fvx.diff_y_T_to_V, "dy"),
    ("diff_y_U_to_X", "u", fvx.diff_y_U_to_X, "dy"), ("diff_x_V_to_X", "v", fvx.diff_x_V_to_X, "dx"),
    ("diff_y_X_to_U", "h", fvx.diff_y_X_to_U, "dy"), ("diff_x_X_to_V", "h", fvx.diff_x_X_to_V, "dx"),
    ("diff_x_U_to_T", "u", fvx.diff_x_U_to_T, "dx"), ("diff_y_V_to_T", "v", fvx.diff_y_V_to_T, "dy"),
]


@pytest.mark.parametrize("variant", ["unmasked", "masked"])
def test_difference2d_goldens(variant):
    need_gpu()
    c = canonical_inputs()
    grid = fx.CartesianGrid2D.from_interior(16, 16, 1.0, 1.0)
    m = fx.Mask2D.from_mask(c["hmask"]) if variant == "masked" else None
    d = fx.Difference2D(grid, m)
    for name, arg, _, _ in DIFF_OPS:
        g = load_golden("Difference2D", name, variant)["out"]
        np.testing.assert_allclose(to_np(getattr(d, name)(c[arg])), g, rtol=1e-12, atol=1e-14)
    for name, args in (("laplacian", ("h",)), ("divergence", ("u", "v")), ("curl", ("u", "v"))):
        g = load_golden("Difference2D", name, variant)["out"]
        out = getattr(d, name)(*[c[a] for a in args])
        np.testing.assert_allclose(to_np(out), g, rtol=1e-12, atol=1e-14)
    gu, gv = d.grad_perp(c["h"])
    g = load_golden("Difference2D", "grad_perp", variant)
    np.testing.assert_allclose(to_np(gu), g["out0"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(to_np(gv), g["out1"], rtol=1e-12, atol=1e-14)
    dv = fx.Divergence2D(grid, m)
    np.testing.assert_allclose(to_np(dv(c["u"], c["v"])),
                               load_golden("Divergence2D", "__call__", variant)["out"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(to_np(dv.noflux(c["u"], c["v"])),
                               load_golden("Divergence2D", "noflux", variant)["out"], rtol=1e-12, atol=1e-14)
    it = fx.Interpolation2D(grid, m)
    np.testing.assert_allclose(to_np(it.T_to_X(c["h"])),
                               load_golden("Interpolation2D", "T_to_X", variant)["out"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(to_np(it.V_to_T(c["v"])),
                               load_golden("Interpolation2D", "V_to_T", variant)["out"], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("variant", ["unmasked", "masked"])
@pytest.mark.parametrize("method", ["upwind1", "weno5"])
def test_advection_goldens(variant, method):
    need_gpu()
    c = canonical_inputs()
    grid = fx.CartesianGrid2D.from_interior(16, 16, 1.0, 1.0)
    m = fx.Mask2D.from_mask(c["hmask"]) if variant == "masked" else None
    out = fx.Advection2D(grid, m)(c["h"], c["u"], c["v"], method=method)
    np.testing.assert_allclose(to_np(out), load_golden("Advection2D", f"__call___{method}", variant)["out"],
                               rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("variant", ["unmasked", "masked"])
def test_diffusion_goldens(variant):
    need_gpu()
    c = canonical_inputs()
    grid = fx.CartesianGrid2D.from_interior(16, 16, 1.0, 1.0)
    m = fx.Mask2D.from_mask(c["hmask"]) if variant == "masked" else None
    np.testing.assert_allclose(to_np(fx.Diffusion2D(grid, m)(c["h"], 1.0)),
                               load_golden("Diffusion2D", "__call__", variant)["out"], rtol=1e-12, atol=1e-14)
    fxx, fyy = fx.Diffusion2D(grid, m).fluxes(c["h"], 1.0)
    g = load_golden("Diffusion2D", "fluxes", variant)
    np.testing.assert_allclose(to_np(fxx), g["out0"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(to_np(fyy), g["out1"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(to_np(fx.BiharmonicDiffusion2D(grid, m)(c["h"], 1e-3)),
                               load_golden("BiharmonicDiffusion2D", "__call__", variant)["out"],
                               rtol=1e-12, atol=1e-14)


ALL_METHODS = ["naive", "upwind1", "upwind2", "upwind3", "weno3", "weno5", "wenoz5",
               "minmod", "van_leer", "superbee", "mc"]


def _fields(ny, nx, seed):
    rng = np.random.default_rng(seed)
    h = rng.standard_normal((ny, nx)).cumsum(1) * 0.1 + rng.standard_normal((ny, nx))
    u = rng.standard_normal((ny, nx))
    v = rng.standard_normal((ny, nx))
    return h, u, v


@pytest.mark.parametrize("method", ALL_METHODS)
@pytest.mark.parametrize("shape", [(5, 5), (37, 70), (130, 67)])
def test_advection_all_methods_unmasked(method, shape):
    need_gpu()
    ny, nx = shape
    h, u, v = _fields(ny, nx, 7)
    dx, dy = 0.3, 0.7
    grid = fx.CartesianGrid2D(nx, ny, 1.0, 1.0, dx, dy)
    ref = fvx.advection_2d(h, u, v, dx, dy, method)
    got = fx.Advection2D(grid)(h, u, v, method=method)
    assert_close_scaled(to_np(got), ref, 1e-12, f"{method} f64")
    got32 = fx.Advection2D(grid)(torch.tensor(h, dtype=torch.float32),
                                 torch.tensor(u, dtype=torch.float32),
                                 torch.tensor(v, dtype=torch.float32), method=method)
    assert_close_scaled(to_np(got32), ref, 1e-5, f"{method} f32")


@pytest.mark.parametrize("method", ALL_METHODS)
def test_advection_all_methods_masked(method):
    need_gpu()
    ny, nx = 71, 90
    h, u, v = _fields(ny, nx, 11)
    hm = random_mask(ny, nx, 3)
    dx, dy = 0.5, 0.25
    grid = fx.CartesianGrid2D(nx, ny, 1.0, 1.0, dx, dy)
    ref = fvx.advection_2d(h, u, v, dx, dy, method, fvx.Mask2D(hm))
    got = fx.Advection2D(grid, fx.Mask2D.from_mask(hm))(h, u, v, method=method)
    assert_close_scaled(to_np(got), ref, 1e-12, f"{method} masked f64")


def test_advection_unknown_method_raises():
    need_gpu()
    grid = fx.CartesianGrid2D.from_interior(8, 8, 1.0, 1.0)
    z = np.zeros((10, 10))
    for bad in ("bogus", "wenoz3"):
        with pytest.raises(ValueError, match="Unknown method"):
            fx.Advection2D(grid)(z, z, z, method=bad)


def test_stencils_random_f32_f64():
    need_gpu()
    ny, nx = 123, 257
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal((2, ny, nx))
    hm = random_mask(ny, nx, 9)
    mk = fvx.Mask2D(hm)
    grid = fx.CartesianGrid2D(nx, ny, 1.0, 1.0, 0.1, 0.2)
    d = fx.Difference2D(grid, fx.Mask2D.from_mask(hm))
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        ta, tb = torch.tensor(a, dtype=dt), torch.tensor(b, dtype=dt)
        assert_close_scaled(to_np(d.laplacian(ta)), fvx.laplacian(a, 0.1, 0.2, mk), tol, "lap")
        assert_close_scaled(to_np(d.divergence(ta, tb)), fvx.divergence(a, b, 0.1, 0.2, mk), tol, "div")
        assert_close_scaled(to_np(d.curl(ta, tb)), fvx.curl(a, b, 0.1, 0.2, mk), tol, "curl")
        bh = fx.BiharmonicDiffusion2D(grid, fx.Mask2D.from_mask(hm))(ta, 2.0)
        assert_close_scaled(to_np(bh), fvx.biharmonic_2d(a, 2.0, 0.1, 0.2, mk), tol, "bih")
        kap = np.abs(b) + 0.1
        df = fx.Diffusion2D(grid, fx.Mask2D.from_mask(hm))(ta, torch.tensor(kap, dtype=dt))
        assert_close_scaled(to_np(df), fvx.diffusion_2d(a, kap, 0.1, 0.2, mk), tol, "diff")


@pytest.mark.parametrize("batch", [(), (3,), (2, 2)])
def test_arakawa_jacobian(batch):
    need_gpu()
    rng = np.random.default_rng(1)
    f = rng.standard_normal((*batch, 34, 67))
    g = rng.standard_normal((*batch, 34, 67))
    ref = fvx.arakawa_jacobian(f, g, 0.3, 0.4)
    got = fx.arakawa_jacobian(f, g, 0.3, 0.4)
    assert got.shape == ref.shape
    assert_close_scaled(to_np(got), ref, 1e-12, "arakawa f64")
    got32 = fx.arakawa_jacobian(torch.tensor(f, dtype=torch.float32),
                                torch.tensor(g, dtype=torch.float32), 0.3, 0.4)
    assert_close_scaled(to_np(got32), ref, 1e-5, "arakawa f32")


def test_periodic_and_rk3():
    need_gpu()
    rng = np.random.default_rng(2)
    a = rng.standard_normal((20, 30))
    for ng in (1, 3):
        np.testing.assert_array_equal(to_np(fx.enforce_periodic(a, ng)), fvx.enforce_periodic(a, ng))
    y = torch.tensor(a)
    out = fx.rk3_ssp_step(y, lambda s: -0.5 * s, 0.1)
    ref = fvx.rk3_ssp_step(a, lambda s: -0.5 * s, 0.1)
    assert_close_scaled(to_np(out), ref, 1e-14, "rk3")


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_lincomb_four_terms_in_place(dtype):
    from finitevolx_amd.timestepping import lincomb

    need_gpu()
    rng = np.random.default_rng(8)
    x, y, z, w = (torch.tensor(v, dtype=dtype, device="cuda") for v in rng.standard_normal((4, 37, 41)))
    x0 = to_np(x)
    want = x0 - 0.75 * to_np(y) + 0.25 * to_np(z) + 0.5 * to_np(w)
    tol = 1e-15 if dtype == torch.float64 else 1e-6
    assert_close_scaled(to_np(lincomb(1.0, x, -0.75, y, 0.25, z, d=0.5, w=w)), want, tol, "lincomb4")
    # in place into the fourth and into the first input
    lincomb(1.0, x, -0.75, y, 0.25, z, out=w, d=0.5, w=w)
    assert_close_scaled(to_np(w), want, tol, "lincomb4 in place")
    lincomb(2.0, x, -1.0, y, out=x)
    assert_close_scaled(to_np(x), 2.0 * x0 - to_np(y), tol, "lincomb3 in place")
    # views one element in: not 16-byte aligned (the scalar kernel)
    xs, ys = x.reshape(-1)[1:], y.reshape(-1)[1:]
    assert_close_scaled(to_np(lincomb(0.5, xs, 3.0, ys)), 0.5 * to_np(xs) + 3.0 * to_np(ys), tol,
                        "lincomb unaligned")


@pytest.mark.parametrize("variant", ["unmasked", "masked"])
def test_momentum_advection_goldens(canon, variant):
    """MomentumAdvection2D (energy scheme) against the reference's goldens."""
    need_gpu()
    from tests.conftest import load_golden

    g = load_golden("MomentumAdvection2D", "__call__", variant)
    m = fx.Mask2D(canon["hmask"]) if variant == "masked" else None
    grid = fx.CartesianGrid2D(Nx=18, Ny=18, Lx=1.0, Ly=1.0, dx=canon["dx"], dy=canon["dy"])
    du, dv = fx.MomentumAdvection2D(grid, m)(canon["u"], canon["v"])
    assert_close_scaled(to_np(du), g["out0"], 1e-12, "du")
    assert_close_scaled(to_np(dv), g["out1"], 1e-12, "dv")


@pytest.mark.parametrize("scheme", ["energy", "enstrophy", "al"])
@pytest.mark.parametrize("masked", [False, True])
def test_momentum_advection_schemes(scheme, masked):
    need_gpu()
    ny, nx = 45, 70
    rng = np.random.default_rng(19)
    u, v = rng.standard_normal((2, ny, nx))
    wet = random_mask(ny, nx, 20) if masked else None
    om = fvx.Mask2D(wet) if masked else None
    ref = fvx.momentum_advection_2d(u, v, 0.3, 0.2, scheme, om)
    grid = fx.CartesianGrid2D(Nx=nx, Ny=ny, Lx=0.3 * (nx - 2), Ly=0.2 * (ny - 2), dx=0.3, dy=0.2)
    madv = fx.MomentumAdvection2D(grid, fx.Mask2D(wet) if masked else None)
    for dtype, tol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        du, dv = madv(torch.tensor(u, dtype=dtype), torch.tensor(v, dtype=dtype), scheme)
        assert_close_scaled(to_np(du), ref[0], tol, f"du {scheme} {dtype}")
        assert_close_scaled(to_np(dv), ref[1], tol, f"dv {scheme} {dtype}")
    with pytest.raises(ValueError):
        madv(torch.tensor(u), torch.tensor(v), "bogus")
