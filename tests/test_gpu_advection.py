"""GPU: the doubly periodic WENO5 tracer step (256^2 interior, ghost width 3,
Lx = Ly = 5.12e6 m, FP64; k^-3 tracer with |k| <= 16, non-divergent velocity
from a streamfunction of the same spectrum scaled to 1 m/s, dt from CFL 0.4)
against the oracle's enforce_periodic + Advection2D + rk3_ssp_step
composition: FP64 to 1e-12, FP32 to 1e-5 (max-norm relative)."""

import numpy as np
import pytest
import torch

import finitevolx_amd as fx
from finitevolx_amd import inputs
from oracle import fvx
from tests.gpu_util import assert_close_scaled, need_gpu, random_mask, to_np

pytestmark = pytest.mark.gpu


def _oracle_step(h, u, v, dx, dy, dt, method, ng, mask=None):
    def rhs(q):
        return fvx.advection_2d(fvx.enforce_periodic(q, ng), u, v, dx, dy, method, mask)
    return fvx.enforce_periodic(fvx.rk3_ssp_step(fvx.enforce_periodic(h, ng), rhs, dt), ng)


def _grid(h, dx):
    ny, nx = h.shape
    return fx.CartesianGrid2D(Nx=nx, Ny=ny, Lx=dx * (nx - 2), Ly=dx * (ny - 2), dx=dx, dy=dx)


def test_config1_weno5_step_matches_oracle():
    need_gpu()
    h, u, v, dx, dt = inputs.advection_config1(256, ng=3)
    assert h.shape == (262, 262)
    ref = _oracle_step(h, u, v, dx, dx, dt, "weno5", 3)
    adv = fx.Advection2D(_grid(h, dx))
    for dtype, tol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        ht = torch.tensor(h, dtype=dtype)
        got = adv.rk3_step(ht, torch.tensor(u, dtype=dtype), torch.tensor(v, dtype=dtype), dt,
                           "weno5", ng=3)
        assert_close_scaled(to_np(got), ref, tol, f"fused step {dtype}")
    # the same step composed from the drop-in operators
    ng = 3
    ut, vt = torch.tensor(u), torch.tensor(v)
    q = fx.rk3_ssp_step(fx.enforce_periodic(torch.tensor(h), ng),
                        lambda s: adv(fx.enforce_periodic(s, ng), ut, vt, "weno5"), dt)
    assert_close_scaled(to_np(fx.enforce_periodic(q, ng)), ref, 1e-12, "composed step")
    # throughput of the fused FP64 step (inputs resident, 200 steps)
    hd = torch.tensor(h, device="cuda")
    ud, vd = torch.tensor(u, device="cuda"), torch.tensor(v, device="cuda")
    for _ in range(5):
        hd = adv.rk3_step(hd, ud, vd, dt, "weno5", ng=3)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(200):
        hd = adv.rk3_step(hd, ud, vd, dt, "weno5", ng=3)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 200
    print(f"[config1] FP64 WENO5 rk3_ssp_step at 256^2: {ms * 1e3:.1f} us/step "
          f"({1e3 / ms:.0f} steps/s, host launches included)", flush=True)


@pytest.mark.parametrize("method", ["upwind1", "upwind3", "weno3", "wenoz5", "van_leer", "mc"])
def test_periodic_step_methods(method):
    need_gpu()
    h, u, v, dx, dt = inputs.advection_config1(48, ng=3, kmax=8.0)
    ref = _oracle_step(h, u, v, dx, dx, dt, method, 3)
    got = fx.Advection2D(_grid(h, dx)).rk3_step(torch.tensor(h), torch.tensor(u),
                                                 torch.tensor(v), dt, method, ng=3)
    assert_close_scaled(to_np(got), ref, 1e-12, method)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_masked_weno5_fp32_and_fp64(dtype, tol):
    """Masked tier dispatch (the 3-layer configuration's tracer is a masked
    FP32 WENO5 field) against the oracle, both precisions."""
    need_gpu()
    n = 70
    rng = np.random.default_rng(11)
    h = rng.standard_normal((n, n))
    u, v = rng.standard_normal((2, n, n))
    wet = random_mask(n, n, 12)
    m = fx.Mask2D(wet)
    ref = fvx.advection_2d(h, u, v, 0.3, 0.4, "weno5", fvx.Mask2D(wet))
    g = fx.CartesianGrid2D(Nx=n, Ny=n, Lx=0.3 * (n - 2), Ly=0.4 * (n - 2), dx=0.3, dy=0.4)
    got = fx.Advection2D(g, m)(torch.tensor(h, dtype=dtype), torch.tensor(u, dtype=dtype),
                               torch.tensor(v, dtype=dtype), "weno5")
    assert_close_scaled(to_np(got), ref, tol, f"masked weno5 {dtype}")


def test_rk3_step_rejects_bad_arguments():
    need_gpu()
    h = torch.zeros(12, 12, dtype=torch.float64)
    adv = fx.Advection2D(fx.CartesianGrid2D.from_interior(10, 10, 1.0, 1.0))
    with pytest.raises(ValueError):
        adv.rk3_step(h, h, h, 0.1, "weno5", ng=5)
    with pytest.raises(ValueError):
        adv.rk3_step(h, h, h, 0.1, "bogus")


RECON = ["naive", "upwind1", "upwind2", "upwind3", "weno3", "wenoz3", "weno5", "wenoz5"]


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_reconstruction2d_face_fluxes(dtype, tol):
    """Every Reconstruction2D.<method>_x / _y (and tvd_x / tvd_y with each
    limiter) on random data with flows of both signs, faces 1..n-2 including
    the wall fallbacks, against the oracle's face fluxes."""
    need_gpu()
    ny, nx = 23, 37
    rng = np.random.default_rng(13)
    h = rng.standard_normal((ny, nx))
    u, v = rng.standard_normal((2, ny, nx))
    rec = fx.Reconstruction2D(fx.CartesianGrid2D.from_interior(nx - 2, ny - 2, 1.0, 1.0))
    t = lambda a: torch.tensor(a, dtype=dtype)  # noqa: E731
    for m in RECON:
        assert_close_scaled(to_np(getattr(rec, f"{m}_x")(t(h), t(u))), fvx.face_flux_x(h, u, m),
                            tol, f"{m}_x")
        assert_close_scaled(to_np(getattr(rec, f"{m}_y")(t(h), t(v))), fvx.face_flux_y(h, v, m),
                            tol, f"{m}_y")
    for lim in ("minmod", "van_leer", "superbee", "mc"):
        assert_close_scaled(to_np(rec.tvd_x(t(h), t(u), lim)), fvx.face_flux_x(h, u, "tvd", lim),
                            tol, f"tvd_x {lim}")
        assert_close_scaled(to_np(rec.tvd_y(t(h), t(v), lim)), fvx.face_flux_y(h, v, "tvd", lim),
                            tol, f"tvd_y {lim}")


def test_reconstruction2d_masked_face_fluxes():
    """The *_masked variants (reconstruction.py:1104-1307) against the
    oracle's tiered upwind_flux on a ragged mask."""
    need_gpu()
    n = 40
    rng = np.random.default_rng(14)
    h = rng.standard_normal((n, n))
    u, v = rng.standard_normal((2, n, n))
    wet = random_mask(n, n, 15)
    om = fvx.Mask2D(wet)
    pm = fx.Mask2D(wet)
    rec = fx.Reconstruction2D(fx.CartesianGrid2D.from_interior(n - 2, n - 2, 1.0, 1.0))
    t = torch.tensor
    for name, meth, lim in (("weno5", "weno5", None), ("wenoz5", "wenoz5", None),
                            ("tvd", "van_leer", "van_leer")):
        funcs = fvx._tier_funcs(meth)
        sizes = tuple(sorted(funcs))
        for dim, axis, vel in ((1, "x", u), (0, "y", v)):
            hier = fvx.narrow_mask_hierarchy(om.get_adaptive_masks(axis, (2, 4, 6)), sizes)
            ref = fvx.upwind_flux(h, vel, dim, funcs, hier)
            f = getattr(rec, f"{name}_{axis}_masked")
            got = f(t(h), t(vel), pm, lim) if lim else f(t(h), t(vel), pm)
            assert_close_scaled(to_np(got), ref, 1e-12, f"{name}_{axis}_masked")


@pytest.mark.parametrize("name", ["euler_step", "heun_step", "rk4_step", "rk3_ssp_step"])
def test_explicit_rk_steps_match_oracle(name):
    """explicit_rk.py steppers on a pytree of device arrays against the oracle
    (the oracle pins each scheme's stability polynomial and order)."""
    need_gpu()
    rng = np.random.default_rng(16)
    a = rng.standard_normal((20, 30))
    b = rng.standard_normal((7, 9))
    f = lambda y: {"a": -0.7 * y["a"] + 0.1 * y["a"] ** 2, "b": 0.3 * y["b"]}  # noqa: E731
    got = getattr(fx, name)({"a": torch.tensor(a), "b": torch.tensor(b)}, f, 0.05)
    ref_a = getattr(fvx, name)(a, lambda y: -0.7 * y + 0.1 * y**2, 0.05)
    ref_b = getattr(fvx, name)(b, lambda y: 0.3 * y, 0.05)
    assert_close_scaled(to_np(got["a"]), ref_a, 1e-14, f"{name} a")
    assert_close_scaled(to_np(got["b"]), ref_b, 1e-14, f"{name} b")


@pytest.mark.parametrize("n", [2048, 4096])
def test_weno5_throughput(n):
    """Advection2D.__call__ (WENO5, one fused tendency pass) at n^2 on random
    fields, FP64 and FP32, unmasked and on the basin mask: FP32 within 1e-5
    of FP64 (max-norm relative), and the time per call (HIP events over 20
    calls) printed with the HBM rate of its algorithmic bytes (read h, u, v,
    write the tendency; + tiers and mask when masked)."""
    need_gpu()
    rng = np.random.default_rng(0)
    fields = [rng.standard_normal((n + 2, n + 2)) for _ in range(3)]
    g = fx.CartesianGrid2D.from_interior(n, n, 1.0, 1.0)
    wet = np.zeros((n + 2, n + 2), dtype=bool)
    wet[1:-1, 1:-1] = inputs.basin_mask(n, n, 3) > 0
    for masked in (False, True):
        adv = fx.Advection2D(g, fx.Mask2D(wet) if masked else None)
        res = {}
        for dt in (torch.float64, torch.float32):
            h, u, v = (torch.tensor(f, dtype=dt, device="cuda") for f in fields)
            out = adv(h, u, v, "weno5")
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(20):
                out = adv(h, u, v, "weno5")
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) / 20 * 1e3
            es = 8 if dt == torch.float64 else 4
            nbytes = n * n * (4 * es + (3 if masked else 0))
            print(f"[weno5] {n}^2 {str(dt)[6:]} masked={masked}: {us:.1f} us/call, "
                  f"{nbytes / us / 1e3:.0f} GB/s ({nbytes / us / 1e3 / 8000:.2f} of HBM peak)",
                  flush=True)
            res[dt] = to_np(out).astype(np.float64)
        assert_close_scaled(res[torch.float32], res[torch.float64], 1e-5, f"fp32 vs fp64 {masked}")
