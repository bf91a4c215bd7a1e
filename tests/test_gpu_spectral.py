"""Spectral Helmholtz solvers (solve_helmholtz_dst/dct/fft).

spectraldiffx v0.0.10 (not vendored) is pinned only through the reference's
closed-form tests (tests/test_solver_wrappers.py:79-92): the DST-I
eigenfunction round trip.  Beyond that each solver is pinned as the exact
inverse of the 5-point operator it diagonalises (oracle operators), FP64 to
1e-10 relative and FP32 to 1e-4.
"""

import numpy as np
import pytest
import torch

import finitevolx_amd as fx
from oracle import fvx
from tests.gpu_util import assert_close_scaled, need_gpu, to_np

pytestmark = pytest.mark.gpu


def test_dst_eigenfunction_roundtrip():
    """Reference tests/test_solver_wrappers.py:79-92."""
    need_gpu()
    Ny, Nx = 10, 12
    dx, dy = 1.0 / (Nx + 1), 1.0 / (Ny + 1)
    j = np.arange(Ny)[:, None]
    i = np.arange(Nx)[None, :]
    psi_exact = np.sin(np.pi * (j + 1) / (Ny + 1)) * np.sin(np.pi * (i + 1) / (Nx + 1))
    zeta = (fx.dst1_eigenvalues(Nx, dx)[0] + fx.dst1_eigenvalues(Ny, dy)[0]) * psi_exact
    psi = fx.streamfunction_from_vorticity(zeta, dx, dy, bc="dst", method="spectral")
    np.testing.assert_allclose(to_np(psi), psi_exact, atol=1e-10)


@pytest.mark.parametrize("shape", [(10, 12), (64, 96), (257, 130)])
@pytest.mark.parametrize("lam", [0.0, 7.0])
def test_dst_inverts_dirichlet_operator(shape, lam):
    need_gpu()
    ny, nx = shape
    dx, dy = 1.0 / nx, 1.3 / ny
    f = np.random.default_rng(0).standard_normal(shape)
    psi = fx.solve_helmholtz_dst(f, dx, dy, lam)
    lev = fvx.build_multigrid(np.ones(shape), dx, dy, lam, n_levels=1, kind="dirichlet")[0]
    assert_close_scaled(fvx.apply_operator(to_np(psi), lev), f, 1e-10, "A psi = f (dst)")
    psi32 = fx.solve_helmholtz_dst(torch.tensor(f, dtype=torch.float32), dx, dy, lam)
    assert_close_scaled(to_np(psi32), to_np(psi), 1e-4, "dst f32")


@pytest.mark.parametrize("shape", [(8, 10), (64, 96), (131, 70)])
def test_dct_inverts_noflux_operator(shape):
    need_gpu()
    ny, nx = shape
    dx, dy = 1.0 / nx, 0.7 / ny
    lam = 3.0
    f = np.random.default_rng(1).standard_normal(shape)
    psi = fx.solve_helmholtz_dct(f, dx, dy, lam)
    lev = fvx.build_multigrid(np.ones(shape), dx, dy, lam, n_levels=1, kind="noflux")[0]
    assert_close_scaled(fvx.apply_operator(to_np(psi), lev), f, 1e-10, "A psi = f (dct)")
    # singular Poisson: zero-mean data is solved up to the constant mode
    f0 = f - f.mean()
    p0 = to_np(fx.solve_poisson_dct(f0, dx, dy))
    lev0 = fvx.build_multigrid(np.ones(shape), dx, dy, 0.0, n_levels=1, kind="noflux")[0]
    assert_close_scaled(fvx.apply_operator(p0, lev0), f0, 1e-9, "poisson dct")
    assert abs(p0.mean()) < 1e-9 * np.abs(p0).max()


@pytest.mark.parametrize("shape", [(8, 10), (64, 96), (33, 50)])
def test_fft_inverts_periodic_operator(shape):
    need_gpu()
    ny, nx = shape
    dx, dy = 2 * np.pi / nx, 2 * np.pi / ny
    f = np.random.default_rng(2).standard_normal(shape)
    psi = fx.solve_helmholtz_fft(f, dx, dy, 2.0)
    ones = np.ones(shape)
    assert_close_scaled(fvx.masked_laplacian(to_np(psi), ones, dx, dy, 2.0), f, 1e-10, "fft")


def test_pv_inversion_spectral_and_batched_layers():
    need_gpu()
    ny, nx = 32, 48
    dx, dy = 1.0 / 33, 1.0 / 49
    q = np.random.default_rng(3).standard_normal((2, 3, ny, nx))
    lam = np.array([0.5, 2.0, 9.0])
    psi = to_np(fx.pv_inversion(q, dx, dy, lam, bc="dst", method="spectral"))
    for b in range(2):
        for k in range(3):
            ref = fvx.dirichlet_helmholtz_direct(q[b, k], np.ones((ny, nx)), dx, dy, lam[k])
            assert_close_scaled(psi[b, k], ref, 1e-10, f"layer {k}")
