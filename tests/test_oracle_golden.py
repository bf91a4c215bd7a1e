"""Pin the oracle to the reference's own golden vectors (rtol 1e-12, atol 1e-14).

Goldens are copied verbatim from the reference's ``tests/fixtures/golden/``
(generated under ``jax_enable_x64``, ``tests/fixtures/_gen_golden.py:53-55``);
inputs restate ``tests/fixtures/inputs.py:141-275``.
"""

import numpy as np
import pytest

from oracle import fvx
from tests.conftest import load_golden

RT, AT = 1e-12, 1e-14


def _check(actual, op, method, variant):
    g = load_golden(op, method, variant)
    if isinstance(actual, tuple):
        for i, a in enumerate(actual):
            np.testing.assert_allclose(a, g[f"out{i}"], rtol=RT, atol=AT)
    else:
        np.testing.assert_allclose(actual, g["out"], rtol=RT, atol=AT)


@pytest.mark.parametrize("variant", ["unmasked", "masked"])
def test_difference2d(canon, variant):
    m = fvx.Mask2D(canon["hmask"]) if variant == "masked" else None
    h, u, v, dx, dy = canon["h"], canon["u"], canon["v"], canon["dx"], canon["dy"]
    _check(fvx.diff_x_T_to_U(h, dx, m), "Difference2D", "diff_x_T_to_U", variant)
    _check(fvx.diff_y_T_to_V(h, dy, m), "Difference2D", "diff_y_T_to_V", variant)
    _check(fvx.diff_y_U_to_X(u, dy, m), "Difference2D", "diff_y_U_to_X", variant)
    _check(fvx.diff_x_V_to_X(v, dx, m), "Difference2D", "diff_x_V_to_X", variant)
    _check(fvx.diff_y_X_to_U(h, dy, m), "Difference2D", "diff_y_X_to_U", variant)
    _check(fvx.diff_x_X_to_V(h, dx, m), "Difference2D", "diff_x_X_to_V", variant)
    _check(fvx.diff_x_U_to_T(u, dx, m), "Difference2D", "diff_x_U_to_T", variant)
    _check(fvx.diff_y_V_to_T(v, dy, m), "Difference2D", "diff_y_V_to_T", variant)
    _check(fvx.laplacian(h, dx, dy, m), "Difference2D", "laplacian", variant)
    _check(fvx.divergence(u, v, dx, dy, m), "Difference2D", "divergence", variant)
    _check(fvx.curl(u, v, dx, dy, m), "Difference2D", "curl", variant)
    _check(fvx.grad_perp(h, dx, dy, m), "Difference2D", "grad_perp", variant)


@pytest.mark.parametrize("variant", ["unmasked", "masked"])
def test_divergence2d(canon, variant):
    m = fvx.Mask2D(canon["hmask"]) if variant == "masked" else None
    u, v, dx, dy = canon["u"], canon["v"], canon["dx"], canon["dy"]
    _check(fvx.divergence(u, v, dx, dy, m), "Divergence2D", "__call__", variant)
    _check(fvx.divergence_noflux(u, v, dx, dy, m), "Divergence2D", "noflux", variant)


@pytest.mark.parametrize("variant", ["unmasked", "masked"])
def test_interpolation2d(canon, variant):
    m = fvx.Mask2D(canon["hmask"]) if variant == "masked" else None
    _check(fvx.T_to_X(canon["h"], m), "Interpolation2D", "T_to_X", variant)
    _check(fvx.V_to_T(canon["v"], m), "Interpolation2D", "V_to_T", variant)


@pytest.mark.parametrize("variant", ["unmasked", "masked"])
@pytest.mark.parametrize("method", ["upwind1", "weno5"])
def test_advection2d(canon, variant, method):
    m = fvx.Mask2D(canon["hmask"]) if variant == "masked" else None
    out = fvx.advection_2d(canon["h"], canon["u"], canon["v"], canon["dx"], canon["dy"], method, m)
    _check(out, "Advection2D", f"__call___{method}", variant)


@pytest.mark.parametrize("variant", ["unmasked", "masked"])
def test_diffusion2d(canon, variant):
    m = fvx.Mask2D(canon["hmask"]) if variant == "masked" else None
    h, dx, dy = canon["h"], canon["dx"], canon["dy"]
    _check(fvx.diffusion_2d(h, 1.0, dx, dy, m), "Diffusion2D", "__call__", variant)
    _check(fvx.diffusion_fluxes(h, 1.0, dx, dy, m), "Diffusion2D", "fluxes", variant)
    _check(fvx.biharmonic_2d(h, 1e-3, dx, dy, m), "BiharmonicDiffusion2D", "__call__", variant)


@pytest.mark.parametrize("variant", ["unmasked", "masked"])
def test_momentum_advection2d(canon, variant):
    m = fvx.Mask2D(canon["hmask"]) if variant == "masked" else None
    _check(fvx.momentum_advection_2d(canon["u"], canon["v"], canon["dx"], canon["dy"], "energy", m),
           "MomentumAdvection2D", "__call__", variant)


def test_planted_mistake_is_caught(canon):
    """A plausible slip (forward instead of backward x-difference) fails the golden."""
    h, dx = canon["u"], canon["dx"]
    wrong = fvx.interior(fvx.diff_x_fwd(h) / dx, h.shape)
    g = load_golden("Difference2D", "diff_x_U_to_T", "unmasked")["out"]
    assert not np.allclose(wrong, g, rtol=RT, atol=AT)
