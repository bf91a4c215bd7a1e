"""Pin the U-Net oracle: hand-written backward vs central finite differences,
pooling/upsampling adjointness, positive homogeneity, Adam closed form."""

import numpy as np
import pytest

from oracle import unet as U


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((5, 6, 3))
    w = rng.standard_normal((3, 3, 2, 3))
    y = U.conv3x3(x, w)
    ref = np.zeros((5, 6, 2))
    for j in range(5):
        for i in range(6):
            for ky in range(3):
                for kx in range(3):
                    jj, ii = j + ky - 1, i + kx - 1
                    if 0 <= jj < 5 and 0 <= ii < 6:
                        ref[j, i] += w[ky, kx] @ x[jj, ii]
    np.testing.assert_allclose(y, ref, atol=1e-12)


def test_pool_up_adjoints():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((8, 6, 2))
    b = rng.standard_normal((4, 3, 2))
    np.testing.assert_allclose(np.sum(U.pool2(a) * b), np.sum(a * U.pool2_bwd(b)), rtol=1e-12)
    np.testing.assert_allclose(np.sum(U.up2(b) * a), np.sum(b * U.up2_bwd(a)), rtol=1e-12)


@pytest.mark.parametrize("channels", [(3, 4), (4, 6, 8)])
def test_backward_matches_finite_differences(channels):
    rng = np.random.default_rng(2)
    Wt = U.init_weights(channels, seed=3)
    r = rng.standard_normal((16, 16))
    gy = rng.standard_normal((16, 16))
    y, cache = U.forward(Wt, r, channels)
    g = U.backward(Wt, cache, gy, channels)
    h = 1e-6
    for name in Wt:
        for _ in range(3):
            idx = tuple(rng.integers(0, s) for s in Wt[name].shape)
            wp = {k: v.copy() for k, v in Wt.items()}
            wm = {k: v.copy() for k, v in Wt.items()}
            wp[name][idx] += h
            wm[name][idx] -= h
            fd = (np.sum(gy * U.forward(wp, r, channels)[0]) - np.sum(gy * U.forward(wm, r, channels)[0])) / (2 * h)
            assert abs(fd - g[name][idx]) <= 1e-5 * max(1.0, abs(fd)), (name, idx, fd, g[name][idx])


def test_positive_homogeneity():
    ch = (4, 6)
    Wt = U.init_weights(ch, seed=5)
    r = np.random.default_rng(6).standard_normal((8, 8))
    y1 = U.forward(Wt, r, ch)[0]
    y2 = U.forward(Wt, 3.5 * r, ch)[0]
    np.testing.assert_allclose(y2, 3.5 * y1, rtol=1e-12, atol=1e-14)


def test_adam_first_step_is_signed_lr():
    Wt = {"a": np.array([1.0, -2.0, 3.0])}
    g = {"a": np.array([0.5, -4.0, 0.0])}
    m = {"a": np.zeros(3)}
    v = {"a": np.zeros(3)}
    U.adam_step(Wt, g, m, v, 1, lr=0.1)
    np.testing.assert_allclose(Wt["a"], [1.0 - 0.1, -2.0 + 0.1, 3.0], atol=1e-7)


def test_bf16_rounding_and_storage_mode():
    # ties to even at 1 + 2^-8 (halfway between 1 and 1 + 2^-7), up above it
    vals = np.array([1.0, 1 + 2.0**-8, 1 + 3 * 2.0**-8, 1 + 2.0**-8 + 2.0**-12, -2.5, 0.0])
    np.testing.assert_array_equal(U.bf16_round(vals),
                                  [1.0, 1.0, 1 + 2.0**-6, 1 + 2.0**-7, -2.5, 0.0])
    ch = (16, 32)
    Wt = U.init_weights(ch, seed=5)
    r = np.random.default_rng(6).standard_normal((16, 16))
    y, _ = U.forward(Wt, r, ch)
    yb, cache = U.forward(Wt, r, ch, bf16=True)
    assert 0 < np.max(np.abs(yb - y)) <= 3e-2 * np.max(np.abs(y))
    # every stored activation is representable in bf16
    for k, v in cache.items():
        for t in (v if isinstance(v, tuple) else (v,)):
            if k != "e0" or t is not v[0]:
                np.testing.assert_array_equal(U.bf16_round(t), t)
