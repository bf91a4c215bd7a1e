"""ctypes binding of ``libfvx_hip.so`` (the C ABI declared in ``include/fvx.h``).

The library is loaded lazily on first use.  There is no fallback: if the
shared library is missing or no GPU is visible, every operator raises.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfvx_hip.so")

_vp = C.c_void_p
_i = C.c_int
_d = C.c_double

# name -> argtypes (restype int unless listed in _RESTYPES)
_SIGS: dict[str, list] = {
    "fvx_abi_version": [],
    "fvx_device_synchronize": [],
}
_RESTYPES = {"fvx_last_error": C.c_char_p}

for _suf in ("f32", "f64"):
    _SIGS.update({
        f"fvx_stencil2d_{_suf}": [_i, _vp, _vp, _vp, _i, _i, _d, _d, _vp, _vp],
        f"fvx_diffusion2d_{_suf}": [_vp, _vp, _d, _vp, _i, _i, _d, _d, _vp, _vp, _vp, _vp],
        f"fvx_diffusion2d_fluxes_{_suf}": [_vp, _vp, _d, _vp, _vp, _i, _i, _d, _d, _vp, _vp, _vp],
        f"fvx_biharmonic2d_{_suf}": [_vp, _vp, _i, _i, _d, _d, _d, _vp, _vp],
        f"fvx_arakawa_jacobian_{_suf}": [_vp, _vp, _vp, _i, _i, _i, _d, _d, _vp],
        f"fvx_masked_laplacian_{_suf}": [_vp, _vp, _vp, _i, _i, _d, _d, _d, _vp],
        f"fvx_enforce_periodic_{_suf}": [_vp, _i, _i, _i, _vp],
        f"fvx_advection2d_{_suf}": [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _vp],
    })

_lib = None


def extra_signatures(sigs: dict[str, list], restypes: dict | None = None) -> None:
    """Register more entry points (used by modules that add ABI groups)."""
    _SIGS.update(sigs)
    if restypes:
        _RESTYPES.update(restypes)
    if _lib is not None:
        _bind(_lib, sigs)


def _bind(lib, sigs):
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)


def load():
    """Load and bind the shared library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -m finitevolx_amd.build` "
            "(finitevolx_amd has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    lib.fvx_last_error.restype = C.c_char_p
    lib.fvx_last_error.argtypes = []
    _bind(lib, _SIGS)
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    """Names this binding expects the library to export."""
    return sorted(set(_SIGS) | {"fvx_last_error"})


class FvxError(RuntimeError):
    pass


def call(name: str, *args) -> int:
    """Call an entry point; raise ``ValueError``/``FvxError`` on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc < 0:
        msg = lib.fvx_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(msg)
        raise FvxError(f"{name}: {msg}")
    return rc


# ---------------------------------------------------------------------------
# Device-array plumbing (torch owns device memory and the stream)
# ---------------------------------------------------------------------------


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("finitevolx_amd needs a ROCm GPU (gfx950); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def as_field(x, dtype=None) -> torch.Tensor:
    """Return ``x`` as a contiguous float32/float64 CUDA tensor."""
    dev = require_gpu()
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(x)))
    if dtype is None:
        dtype = t.dtype if t.dtype in (torch.float32, torch.float64) else torch.float64
    return t.to(device=dev, dtype=dtype).contiguous()


def as_mask(x) -> torch.Tensor | None:
    """Boolean-like array -> contiguous uint8 CUDA tensor (None stays None)."""
    if x is None:
        return None
    dev = require_gpu()
    if isinstance(x, torch.Tensor):
        return (x != 0).to(device=dev, dtype=torch.uint8).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x) != 0).astype(np.uint8)).to(dev)


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def suffix(t: torch.Tensor) -> str:
    if t.dtype == torch.float64:
        return "f64"
    if t.dtype == torch.float32:
        return "f32"
    raise TypeError(f"unsupported dtype {t.dtype}; use float32 or float64")
