"""Learned U-Net preconditioner (no reference counterpart).

A bias-free convolutional U-Net (BF16 weights and activations, FP32
accumulation on matrix cores) trained on (residual, error) pairs that the
multigrid solver generates; used inside flexible PCG where the FP64 residual
certifies convergence and plain multigrid-PCG is the fallback.  Forward,
backward and Adam run in ``libfvx_hip.so``; this module orchestrates data
and training.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from finitevolx_amd import _lib

_vp, _i, _d, _i64 = _lib._vp, _lib._i, _lib._d, C.c_int64
_lib.extra_signatures({
    "fvx_unet_create": [_i, _vp, _i, _i, C.POINTER(_vp)],
    "fvx_unet_destroy": [_vp],
    "fvx_unet_weights": [_vp, _vp, _i, _vp],
    "fvx_unet_forward_raw": [_vp, _vp, _vp, _i, _i, _vp],
    "fvx_unet_backward_raw": [_vp, _vp, _vp, _i, _i, _vp],
    "fvx_unet_grads": [_vp, _vp, _i, _vp],
    "fvx_unet_train_step": [_vp, _vp, _vp, _i, _i, _i, _d, C.POINTER(C.c_double), _vp],
    "fvx_unet_apply_f32": [_vp, _vp, _i64, _vp, _i64, _i, _i, _d, _vp, _vp],
    "fvx_conv3x3_bf16": [_vp, _i, _vp, _i, _vp, _i, _i, _i, _i, _vp, _vp],
}, {"fvx_unet_num_weights": _i64})
_lib._SIGS["fvx_unet_num_weights"] = [_vp]


def weight_shapes(channels):
    """Per-layer fp32 weight shapes (3, 3, Cout, Cin) in the library's order."""
    L = len(channels)
    shapes = []
    cin = 1
    for c in channels:
        shapes.append((3, 3, c, cin))
        shapes.append((3, 3, c, c))
        cin = c
    for l in range(L - 2, -1, -1):
        shapes.append((3, 3, channels[l], channels[l + 1]))
        shapes.append((3, 3, channels[l], channels[l]))
    shapes.append((3, 3, 1, channels[0]))
    return shapes


def layer_names(channels):
    L = len(channels)
    names = []
    for l in range(L):
        names += [f"e{l}a", f"e{l}b"]
    for l in range(L - 2, -1, -1):
        names += [f"u{l}", f"d{l}"]
    return names + ["out"]


def to_bf16(x: np.ndarray) -> np.ndarray:
    """Round float64/32 to the nearest bfloat16 (ties to even), as float32."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


class UNet:
    """Device U-Net plan sized for fields up to ``(H, W)``."""

    def __init__(self, channels=(16, 32, 48, 64), H: int = 256, W: int = 256, seed: int = 0,
                 weights: dict | None = None):
        _lib.require_gpu()
        self.channels = tuple(int(c) for c in channels)
        self.H, self.W = H, W
        arr = (C.c_int * len(self.channels))(*self.channels)
        h = _vp()
        _lib.call("fvx_unet_create", len(self.channels), arr, H, W, C.byref(h))
        self._h = h
        self.shapes = weight_shapes(self.channels)
        self.names = layer_names(self.channels)
        if weights is None:
            rng = np.random.default_rng(seed)
            weights = {n: rng.standard_normal(s) * np.sqrt(2.0 / (9 * s[3]))
                       for n, s in zip(self.names, self.shapes)}
        self.set_weights(weights)

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _lib.call("fvx_unet_destroy", self._h)
            except Exception:
                pass
            self._h = None

    # --- weights -----------------------------------------------------------
    def set_weights(self, weights: dict):
        flat = np.concatenate([np.asarray(weights[n], np.float32).ravel() for n in self.names])
        t = torch.from_numpy(flat).to(_lib.require_gpu())
        _lib.call("fvx_unet_weights", self._h, t.data_ptr(), 0, _lib.stream())
        torch.cuda.synchronize()

    def get_weights(self) -> dict:
        n = sum(int(np.prod(s)) for s in self.shapes)
        t = torch.empty(n, dtype=torch.float32, device=_lib.require_gpu())
        _lib.call("fvx_unet_weights", self._h, t.data_ptr(), 1, _lib.stream())
        return self._unflatten(t.cpu().numpy())

    def _unflatten(self, flat):
        out, o = {}, 0
        for n, s in zip(self.names, self.shapes):
            k = int(np.prod(s))
            out[n] = flat[o:o + k].reshape(s).astype(np.float64)
            o += k
        return out

    # --- raw forward / backward (network units) ----------------------------
    def forward(self, x):
        x = _lib.as_field(x, torch.float32)
        y = torch.empty_like(x)
        _lib.call("fvx_unet_forward_raw", self._h, x.data_ptr(), y.data_ptr(), x.shape[0],
                  x.shape[1], _lib.stream())
        return y

    def grads(self, x, gy) -> dict:
        """Weight gradients of sum(gy * net(x))."""
        x = _lib.as_field(x, torch.float32)
        gy = _lib.as_field(gy, torch.float32)
        _lib.call("fvx_unet_backward_raw", self._h, x.data_ptr(), gy.data_ptr(), x.shape[0],
                  x.shape[1], _lib.stream())
        n = sum(int(np.prod(s)) for s in self.shapes)
        t = torch.empty(n, dtype=torch.float32, device=x.device)
        _lib.call("fvx_unet_grads", self._h, t.data_ptr(), 1, _lib.stream())
        return self._unflatten(t.cpu().numpy())

    def train_step(self, xb: torch.Tensor, eb: torch.Tensor, lr: float) -> float:
        """Adam step on a batch (B, H, W) of inputs and targets; returns the loss."""
        xb = xb.to(torch.float32).contiguous()
        eb = eb.to(torch.float32).contiguous()
        loss = C.c_double()
        _lib.call("fvx_unet_train_step", self._h, xb.data_ptr(), eb.data_ptr(), xb.shape[0],
                  xb.shape[1], xb.shape[2], float(lr), C.byref(loss), _lib.stream())
        return loss.value

    # --- preconditioner -----------------------------------------------------
    def apply(self, r: torch.Tensor, out_scale: float, mask: torch.Tensor | None = None):
        """z = out_scale * rms(r) * net(r / rms(r)) * mask on an FP32 field."""
        r = _lib.as_field(r, torch.float32)
        z = torch.empty_like(r)
        _lib.call("fvx_unet_apply_f32", self._h, r.data_ptr(), r.shape[1], z.data_ptr(),
                  z.shape[1], r.shape[0], r.shape[1], float(out_scale), _lib.ptr(mask),
                  _lib.stream())
        return z


def conv3x3_bf16(x: torch.Tensor, w: np.ndarray, ups: bool = False, relu: bool = False):
    """One matrix-core 3x3 conv: x (H, W, Cin) bf16, w (3, 3, Cout, Cin) -> (H', W', Cout)."""
    x = x.contiguous()
    Hx, Wx, cin = x.shape
    cout = w.shape[2]
    H, W = (2 * Hx, 2 * Wx) if ups else (Hx, Wx)
    wt = torch.from_numpy(np.ascontiguousarray(w, dtype=np.float32)).to(x.device)
    out = torch.empty((H, W, cout), dtype=torch.bfloat16, device=x.device)
    scratch = torch.empty(9 * cout * cin, dtype=torch.bfloat16, device=x.device)
    _lib.call("fvx_conv3x3_bf16", x.data_ptr(), cin, wt.data_ptr(), cout, out.data_ptr(), H, W,
              int(ups), int(relu), scratch.data_ptr(), _lib.stream())
    return out


# ---------------------------------------------------------------------------
# Training data from the multigrid solver, and training
# ---------------------------------------------------------------------------


def generate_pairs(mg, rhs_fields, vcycles=(1, 2, 3), sweeps: int = 2, rtol=1e-11):
    """Coarse-level (residual, error) training pairs generated by the
    multigrid solver, for the network's place in the learned two-grid cycle.

    For each right-hand side b: the converged solution x* (multigrid-PCG to
    ``rtol``) and the iterates x_k of k plain V-cycles from zero (k in
    ``vcycles``, plus x_0 = 0) give residual/error pairs (r_k = b - A x_k,
    e_k = x* - x_k).  Each pair is taken through the cycle's pre-smoothing,
    z = S^sweeps(0; r_k), res = r_k - A z, e = e_k - z, and restricted to
    level 1 with the solver's own restriction: (R res, R e).  ``mg`` is an
    FP64 ``MultigridSolver`` with the same operator and omega as the solve.
    Returns FP32 device tensors (coarse residuals, coarse errors) and the
    output scale (median rms(error) / rms(residual)) to train and apply with.
    """
    from finitevolx_amd.solvers import apply_operator, restrict, weighted_jacobi

    if mg.n_levels < 2:
        raise ValueError("the learned two-grid cycle needs at least two multigrid levels")
    lev0 = mg.levels[0]
    rs, es, ratios = [], [], []
    for b in rhs_fields:
        b = _lib.as_field(b, torch.float64)
        xs, info = mg.pcg(b, rtol=rtol, atol=0.0, max_steps=400)
        if not info.converged:
            raise RuntimeError(f"pair generation solve did not converge: {info}")
        iterates = [torch.zeros_like(b)]
        u = None
        for k in range(1, max(vcycles) + 1):
            u = mg.v_cycle(u, b)
            if k in vcycles:
                iterates.append(u)
        for x in iterates:
            r = b - apply_operator(x, lev0)
            e = xs - x
            if sweeps > 0:
                z = weighted_jacobi(None, r, lev0, mg.omega, sweeps)
                e = e - z
                r = r - apply_operator(z, lev0)
            rc = restrict(r, lev0)
            ec = restrict(e, lev0)
            rn = float(torch.sqrt(torch.mean(rc * rc)))
            if rn == 0.0:
                continue
            ratios.append(float(torch.sqrt(torch.mean(ec * ec))) / rn)
            rs.append(rc.float())
            es.append(ec.float())
    return rs, es, float(np.median(ratios))


def crop_batch(rs, es, crop: int, B: int, out_scale: float, gen: torch.Generator,
               min_signal: float = 0.25, tries: int = 16):
    """Random ``crop x crop`` windows, normalised per crop: x = r / rms(r),
    target = e / (rms(r) * out_scale).  Windows whose residual rms is below
    ``min_signal`` times the whole field's (mostly land, or a quiet corner:
    their normalised targets are outliers that destabilise training) are
    redrawn, up to ``tries`` times."""
    xb, eb = [], []
    n = len(rs)
    for _ in range(B):
        for _t in range(tries):
            k = int(torch.randint(n, (1,), generator=gen))
            H, W = rs[k].shape
            j0 = int(torch.randint(H - crop + 1, (1,), generator=gen))
            i0 = int(torch.randint(W - crop + 1, (1,), generator=gen))
            r = rs[k][j0:j0 + crop, i0:i0 + crop]
            e = es[k][j0:j0 + crop, i0:i0 + crop]
            s = torch.sqrt(torch.mean(r * r)).clamp_min(1e-30)
            if float(s) >= min_signal * float(torch.sqrt(torch.mean(rs[k] * rs[k]))):
                break
        xb.append(r / s)
        eb.append(e / (s * out_scale))
    return torch.stack(xb).contiguous(), torch.stack(eb).contiguous()


def train(net: UNet, rs, es, out_scale: float, steps: int = 600, crop: int = 128, batch: int = 8,
          lr: float = 2e-3, lr_final: float = 2e-4, seed: int = 0, log_every: int = 0):
    """Adam training on MG-generated pairs; returns the loss history."""
    gen = torch.Generator().manual_seed(seed)
    hist = []
    for it in range(steps):
        frac = it / max(1, steps - 1)
        lr_it = lr * (lr_final / lr) ** frac
        xb, eb = crop_batch(rs, es, crop, batch, out_scale, gen)
        loss = net.train_step(xb, eb, lr_it)
        hist.append(loss)
        if log_every and it % log_every == 0:
            print(f"[train] step {it} loss {loss:.4e}", flush=True)
    return hist


def fit_preconditioner(mask, dx: float, dy: float, lambda_: float, rhs_fields, *,
                       coeff=None, kind: str = "dirichlet", n_pre: int = 2, n_post: int = 2,
                       n_coarse: int = 20, omega: float = 0.8, sweeps: int = 2,
                       channels=(16, 32, 48, 64), steps: int = 600, crop: int = 128,
                       batch: int = 8, lr: float = 2e-3, lr_final: float = 2e-4, seed: int = 0,
                       vcycles=(1, 2, 3)):
    """Train a U-Net preconditioner for one Helmholtz operator on the GPU.

    Builds the FP64 multigrid solver of the operator (``coeff``: the
    variable coefficient of the no-flux operator, or None), generates coarse
    (residual, error) pairs from it on ``rhs_fields`` and trains the network
    (sized for multigrid level 1) with this build's kernels.  Returns
    ``(net, out_scale, loss_history)``; pass the network, ``out_scale`` and
    the same ``sweeps`` to ``MultigridSolver.precond_solve`` or
    ``QGModel.set_preconditioner``.
    """
    from finitevolx_amd.solvers import build_multigrid_solver

    m = np.asarray(mask, dtype=np.float64)
    mg = build_multigrid_solver(m, dx, dy, lambda_, coeff, kind=kind, n_pre=n_pre,
                                n_post=n_post, n_coarse=n_coarse, omega=omega)
    rs, es, out_scale = generate_pairs(mg, [np.asarray(f) * m for f in rhs_fields],
                                       vcycles=vcycles, sweeps=sweeps)
    del mg
    H, W = rs[0].shape
    net = UNet(channels, H, W, seed=seed)
    crop = min(crop, H, W)
    hist = train(net, rs, es, out_scale, steps=steps, crop=crop, batch=batch, lr=lr,
                 lr_final=lr_final, seed=seed)
    return net, out_scale, hist
