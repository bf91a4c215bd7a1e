"""GPU: the allocation-free periodic advection step (stage arrays in the
library's static device scratch) against the allocating entry point, through
the C ABI: bit-identical on two calls in a row, the second smaller; sizes
beyond the scratch are refused with an error before any device work."""

import numpy as np
import pytest
import torch

from finitevolx_amd import _lib, inputs
from finitevolx_amd.advection import method_codes
from tests.gpu_util import need_gpu

pytestmark = pytest.mark.gpu


def _call(entry, h, u, v, dx, dt, method="weno5", ng=3):
    m, lim = method_codes(method)
    ny, nx = h.shape
    return _lib.call(f"{entry}_{_lib.suffix(h)}", m, lim, h.data_ptr(), u.data_ptr(), v.data_ptr(),
                     None, None, None, ny, nx, ng, float(dx), float(dx), float(dt), _lib.stream())


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_step_matches_allocating_step(dtype):
    need_gpu()
    for n in (256, 96):  # the second call is smaller
        h, u, v, dx, dt = inputs.advection_config1(n, ng=3)
        dev = torch.device("cuda")
        u_t = torch.tensor(u, dtype=dtype, device=dev)
        v_t = torch.tensor(v, dtype=dtype, device=dev)
        a = torch.tensor(h, dtype=dtype, device=dev)
        b = a.clone()
        _call("fvx_advection_rk3_step", a, u_t, v_t, dx, dt)
        _call("fvx_advection_rk3_step_noalloc", b, u_t, v_t, dx, dt)
        torch.cuda.synchronize()
        assert torch.equal(a, b), f"n={n}: max diff {(a - b).abs().max().item()}"
        assert torch.isfinite(b).all()


def test_noalloc_step_refuses_sizes_beyond_scratch():
    need_gpu()
    # 2 x 2100^2 FP64 stage arrays (70.6 MB) exceed the 32 MiB scratch; the
    # entry point checks before launching, so these arrays are never read.
    n = 2100
    big = torch.zeros(n, n, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="scratch"):
        _call("fvx_advection_rk3_step_noalloc", big, big, big, 1.0, 0.1)
