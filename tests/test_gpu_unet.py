"""Learned preconditioner kernels against the U-Net oracle (oracle/unet.py).

Single conv: bf16-exact inputs and weights, FP32 accumulation -> within one
bf16 ulp of the oracle (rounded to bf16).  Whole network: against the oracle
run at the same storage precision (``bf16=True``: activations and activation
gradients rounded to bfloat16 where the device stores them; matrix-core
layers with bf16 weights, first/last layers FP32): output within 1e-2
(max-norm relative) and weight gradients within 1e-1 (Frobenius relative)
and 2e-1 (max-norm).  Measured: <= 8e-3 Frobenius for the 2-level network,
up to 6.3e-2 Frobenius / 1.0e-1 max for the deepest decoder layer of the
4-level one -- FP32 accumulation order flips stored values by one bf16 ulp
and ReLU gates near zero, and those flips compound with depth.  An indexing
error gives O(1) differences.  Against the exact FP64 network the same bf16
storage alone moves deep-layer gradients by up to ~20 %, so the exact
comparison is not a useful gate.
"""

import numpy as np
import pytest
import torch

from finitevolx_amd import learned as Lm
from oracle import unet as U
from tests.gpu_util import need_gpu

pytestmark = pytest.mark.gpu


def _bf16(t):
    return torch.from_numpy(Lm.to_bf16(t)).to("cuda", torch.bfloat16)


@pytest.mark.parametrize("cin,cout,ups", [(16, 16, False), (32, 48, False), (64, 16, True),
                                          (48, 64, False), (16, 32, True)])
def test_single_conv_mfma(cin, cout, ups):
    need_gpu()
    rng = np.random.default_rng(0)
    Hx, Wx = (12, 40) if ups else (20, 70)
    x = Lm.to_bf16(rng.standard_normal((Hx, Wx, cin))).astype(np.float64)
    w = Lm.to_bf16(rng.standard_normal((3, 3, cout, cin)) * 0.2).astype(np.float64)
    xin = U.up2(x) if ups else x
    ref = U.conv3x3(xin, w)
    got = Lm.conv3x3_bf16(_bf16(x), w, ups=ups).float().cpu().numpy()
    refb = Lm.to_bf16(ref)
    ulp = np.abs(refb).astype(np.float64) * 2.0**-7 + 1e-30
    assert np.all(np.abs(got - refb) <= ulp + 1e-6 * np.abs(ref).max()), np.max(np.abs(got - refb) / ulp)
    got_r = Lm.conv3x3_bf16(_bf16(x), w, ups=ups, relu=True).float().cpu().numpy()
    assert np.all(got_r >= 0) and np.allclose(got_r, np.maximum(got, 0))


@pytest.mark.parametrize("channels", [(16, 32), (16, 32, 48, 64)])
def test_forward_and_gradients_match_oracle(channels):
    need_gpu()
    H = W = 64
    net = Lm.UNet(channels, H, W, seed=1)
    # matrix-core layers run on bf16 weights; the 1-channel edge layers on FP32
    Wt = {k: (v if k in ("e0a", "out") else Lm.to_bf16(v).astype(np.float64))
          for k, v in net.get_weights().items()}
    rng = np.random.default_rng(2)
    x = Lm.to_bf16(rng.standard_normal((H, W))).astype(np.float64)
    ref, cache = U.forward(Wt, x, channels, bf16=True)
    got = net.forward(torch.tensor(x, dtype=torch.float32)).cpu().numpy()
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err < 1e-2, err
    gy = rng.standard_normal((H, W))
    gref = U.backward(Wt, cache, gy, channels, bf16=True)
    g = net.grads(torch.tensor(x, dtype=torch.float32), torch.tensor(gy, dtype=torch.float32))
    for k in gref:
        e = np.max(np.abs(g[k] - gref[k])) / max(np.max(np.abs(gref[k])), 1e-30)
        f = np.linalg.norm(g[k] - gref[k]) / max(np.linalg.norm(gref[k]), 1e-30)
        print(f"[unet] {channels} {k}: max {e:.3e} fro {f:.3e}", flush=True)
        assert e < 2e-1 and f < 1e-1, (k, e, f)


def test_training_reduces_loss():
    need_gpu()
    H = W = 32
    net = Lm.UNet((16, 32), H, W, seed=3)
    rng = np.random.default_rng(4)
    xb = torch.tensor(rng.standard_normal((4, H, W)), dtype=torch.float32, device="cuda")
    eb = xb * 0.5  # a learnable linear target
    l0 = net.train_step(xb, eb, 1e-3)
    for _ in range(60):
        l1 = net.train_step(xb, eb, 1e-3)
    assert l1 < 0.5 * l0, (l0, l1)
