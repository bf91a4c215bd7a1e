"""Helpers shared by the GPU parity tests."""

import numpy as np
import pytest
import torch


def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


def to_np(t):
    return t.detach().cpu().numpy().astype(np.float64)


def assert_close_scaled(actual, expected, rel, what=""):
    """max|a - e| <= rel * max(max|e|, tiny): the norm-relative parity bar."""
    a = np.asarray(actual, np.float64)
    e = np.asarray(expected, np.float64)
    assert a.shape == e.shape, f"{what}: shape {a.shape} vs {e.shape}"
    scale = max(np.max(np.abs(e)), 1e-300)
    err = np.max(np.abs(a - e)) if a.size else 0.0
    assert err <= rel * scale, f"{what}: max err {err:.3e} > {rel:.1e} * {scale:.3e}"


def random_mask(ny, nx, seed, land_frac=0.25):
    """Ragged land mask: smooth random field thresholded, ghost ring dry."""
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((ny, nx))
    for _ in range(3):
        f = (f + np.roll(f, 1, 0) + np.roll(f, -1, 0) + np.roll(f, 1, 1) + np.roll(f, -1, 1)) / 5
    m = f > np.quantile(f, land_frac)
    m[0, :] = m[-1, :] = m[:, 0] = m[:, -1] = False
    return m
