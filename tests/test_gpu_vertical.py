"""Vertical modes: closed forms (reference tests/test_vertical_modes.py) on the
host and the GPU layer <-> mode transform against the oracle."""

import numpy as np
import pytest
import torch

import finitevolx_amd as fx
from oracle import fvx
from tests.gpu_util import assert_close_scaled, need_gpu, to_np


def test_coupling_matrix_and_modes_host():
    H = [400.0, 1100.0, 2600.0]
    g = [9.81, 0.025, 0.0125]
    A = fx.build_coupling_matrix(H, g)
    np.testing.assert_allclose(A, fvx.build_coupling_matrix(H, g), rtol=1e-15)
    assert A[0, 1] == pytest.approx(-1.0 / (400.0 * 0.025))
    r, l2m, m2l = fx.decompose_vertical_modes(A, 9.375e-5)
    np.testing.assert_allclose(l2m @ m2l, np.eye(3), atol=1e-12)
    ro, _, _ = fvx.decompose_vertical_modes(A, 9.375e-5)
    np.testing.assert_allclose(r, ro, rtol=1e-12)
    # single layer: Rd = sqrt(H g') / |f0|  (reference tests/test_vertical_modes.py)
    r1, _, _ = fx.decompose_vertical_modes(fx.build_coupling_matrix([1000.0], [9.81]), 1e-4)
    assert r1[0] == pytest.approx(np.sqrt(1000.0 * 9.81) / 1e-4, rel=1e-12)


@pytest.mark.gpu
def test_layer_mode_roundtrip_gpu():
    need_gpu()
    A = fx.build_coupling_matrix([400.0, 1100.0, 2600.0], [9.81, 0.025, 0.0125])
    _, l2m, m2l = fx.decompose_vertical_modes(A, 9.375e-5)
    f = np.random.default_rng(0).standard_normal((3, 70, 90))
    for dt, tol in ((torch.float64, 1e-14), (torch.float32, 1e-6)):
        modes = fx.layer_to_mode(torch.tensor(f, dtype=dt), l2m)
        assert_close_scaled(to_np(modes), fvx.layer_to_mode(f, l2m), tol, "l2m")
        back = fx.mode_to_layer(modes, m2l)
        assert_close_scaled(to_np(back), f, 10 * tol, "roundtrip")
