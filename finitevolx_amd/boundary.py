"""Ghost-ring fills used between RK stages (``boundary/boundary.py:17-78``)."""

from __future__ import annotations

import torch

from finitevolx_amd import _lib


def enforce_periodic(field, ng: int = 1):
    """Periodic ghost fill of width ``ng`` (returns a new array)."""
    f = _lib.as_field(field).clone()
    ny, nx = f.shape
    _lib.call(f"fvx_enforce_periodic_{_lib.suffix(f)}", f.data_ptr(), ny, nx, int(ng),
              _lib.stream())
    return f


def pad_interior(field, mode: str = "constant"):
    """Strip the ghost ring and re-pad it (``boundary.py:17-49``); zero or edge."""
    f = _lib.as_field(field)
    inner = f[1:-1, 1:-1]
    if mode == "constant":
        return torch.nn.functional.pad(inner[None, None], (1, 1, 1, 1))[0, 0].contiguous()
    if mode == "edge":
        return torch.nn.functional.pad(inner[None, None], (1, 1, 1, 1), mode="replicate")[0, 0].contiguous()
    raise ValueError(f"unsupported pad mode {mode!r}")
