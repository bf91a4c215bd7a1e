"""Multi-layer QG with a passive tracer on the GPU against the oracle.

The 3-layer stratification of the 4096^2 configuration (H = [400, 1100,
2600] m, g' = [9.81, 0.025, 0.0125] m/s^2): at small sizes every step is
compared with the oracle's direct-solve composition; at 4096^2 the run is
checked through size-independent properties (every modal solve certified,
psi and tracer zero on land, the flux-form tracer total conserved).
"""

import time

import numpy as np
import pytest
import torch

from oracle import fvx as O

pytestmark = pytest.mark.gpu

H = [400.0, 1100.0, 2600.0]
G = [9.81, 0.025, 0.0125]


def _fields(n, L, seed):
    from finitevolx_amd import multilayer

    return multilayer.config5_setup(n, L=L, seed=seed)


@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_geostrophic_faces_match_oracle(masked, dtype):
    from finitevolx_amd import _lib
    from finitevolx_amd.multilayer import MultiLayerQG  # noqa: F401  (registers signatures)

    n = 70
    rng = np.random.default_rng(3)
    psi = np.zeros((n + 2, n + 2))
    psi[1:-1, 1:-1] = rng.standard_normal((n, n))
    mask = None
    if masked:
        from finitevolx_amd import inputs

        wet = np.zeros((n + 2, n + 2), bool)
        wet[1:-1, 1:-1] = inputs.basin_mask(n, n, 11)
        psi *= wet
        mask = O.Mask2D(wet)
    p = O.QGParams(n, n, 7e5, 8.4e5)
    u0, v0 = O.tracer_velocity(psi, p, mask)
    dev = torch.device("cuda")
    ps = torch.tensor(psi, device=dev)
    u = torch.zeros(n + 2, n + 2, dtype=dtype, device=dev)
    v = torch.zeros_like(u)
    m = None
    if masked:
        from finitevolx_amd.mask import Mask2D

        m = Mask2D(wet)
    mp = [None if m is None else m.device(k) for k in ("xy_corner_strict", "u", "v")]
    suf = "f64" if dtype == torch.float64 else "f32"
    _lib.call(f"fvx_geostrophic_faces_{suf}", ps.data_ptr(), _lib.ptr(mp[0]), _lib.ptr(mp[1]),
              _lib.ptr(mp[2]), u.data_ptr(), v.data_ptr(), n + 2, n + 2, p.dx, p.dy,
              _lib.stream())
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    scale = max(np.abs(u0).max(), np.abs(v0).max())
    np.testing.assert_allclose(u.double().cpu().numpy(), u0, rtol=0, atol=tol * scale)
    np.testing.assert_allclose(v.double().cpu().numpy(), v0, rtol=0, atol=tol * scale)


@pytest.mark.parametrize("masked", [False, True])
def test_advection_stage_matches_composition(masked):
    from finitevolx_amd import _lib
    from finitevolx_amd.advection import method_codes
    from finitevolx_amd.mask import Mask2D
    from finitevolx_amd.multilayer import MultiLayerQG  # noqa: F401

    n = 90
    rng = np.random.default_rng(4)
    hs, hn, u, v = rng.standard_normal((4, n, n))
    mask = mm = None
    if masked:
        from finitevolx_amd import inputs

        wet = np.zeros((n, n), bool)
        wet[2:-2, 2:-2] = inputs.basin_mask(n - 4, n - 4, 12)
        mask, mm = O.Mask2D(wet), Mask2D(wet)
        u, v = u * mask.u, v * mask.v
    dx, dy, a, b, cdt = 0.7, 0.9, 0.75, 0.25, 0.3
    want = a * hn + b * hs + cdt * O.advection_2d(hs, u, v, dx, dy, "weno5", mask)
    dev = torch.device("cuda")
    t = [torch.tensor(x, device=dev) for x in (hs, u, v, hn)]
    out = torch.zeros_like(t[0])
    m, lim = method_codes("weno5")
    tiers = (None, None) if mm is None else (mm.device("tier_x"), mm.device("tier_y"))
    _lib.call("fvx_advection_stage_f64", m, lim, t[0].data_ptr(), t[1].data_ptr(),
              t[2].data_ptr(), t[3].data_ptr(), out.data_ptr(), _lib.ptr(tiers[0]),
              _lib.ptr(tiers[1]), _lib.ptr(None if mm is None else mm.device("h")), n, n, dx, dy,
              a, b, cdt, _lib.stream())
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=1e-12, atol=1e-12)
    # out may alias hn (the last stage writes into the step's start state)
    _lib.call("fvx_advection_stage_f64", m, lim, t[0].data_ptr(), t[1].data_ptr(),
              t[2].data_ptr(), t[3].data_ptr(), t[3].data_ptr(), _lib.ptr(tiers[0]),
              _lib.ptr(tiers[1]), _lib.ptr(None if mm is None else mm.device("h")), n, n, dx, dy,
              a, b, cdt, _lib.stream())
    np.testing.assert_allclose(t[3].cpu().numpy(), want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("masked", [True, False])
def test_multilayer_matches_oracle(masked):
    from finitevolx_amd.multilayer import MultiLayerQG

    n, L = 64, 1.6e6  # 25 km cells: the modal radii (42 km, 25 km) stay resolved
    kw = _fields(n, L, seed=1)
    kw["F0"] = 2e-12  # wind forcing on the top layer
    if not masked:
        kw["mask"] = None
    model = MultiLayerQG(**kw, tracer_dtype=torch.float64, dt=900.0)
    wet = (np.ones((n, n)) if kw["mask"] is None else np.asarray(kw["mask"], float))
    p = O.MultiLayerParams(n, n, L, L, H, G, dt=900.0, F0=2e-12)
    F = p.forcing()
    fullwet = np.zeros((n + 2, n + 2), bool)
    fullwet[1:-1, 1:-1] = wet > 0
    mask = O.Mask2D(fullwet) if kw["mask"] is not None else None
    q = np.zeros((3, n + 2, n + 2))
    q[:, 1:-1, 1:-1] = kw["q0"] * wet
    c = np.zeros((n + 2, n + 2))
    c[1:-1, 1:-1] = kw["c0"] * wet
    for step in range(8):
        q, c, psi = O.multilayer_step(q, c, p, F, wet, mask)
        info = model.step(1)
        assert info["converged"], (step, info)
        torch.cuda.synchronize()
        gq = model.q.cpu().numpy()
        gc = model.c.cpu().numpy()
        print(f"step {step}: |dq|/|q| {np.abs(gq - q).max() / np.abs(q).max():.2e} "
              f"|dc| {np.abs(gc - c).max():.2e} budgets {model.budgets}", flush=True)
        np.testing.assert_allclose(gq, q, rtol=0, atol=1e-6 * np.abs(q).max())
        np.testing.assert_allclose(gc, c, rtol=0, atol=1e-7 * np.abs(c).max())
    gp = model.psi.cpu().numpy()
    np.testing.assert_allclose(gp, psi, rtol=0, atol=1e-6 * np.abs(psi).max())
    assert model.steps_done == 8 and model.recaptures >= 1


def test_multilayer_4096_config():
    from finitevolx_amd.multilayer import MultiLayerQG

    t0 = time.perf_counter()
    kw = _fields(4096, 1.024e7, seed=0)
    model = MultiLayerQG(**kw)
    t_setup = time.perf_counter() - t0
    wet = torch.tensor(np.asarray(kw["mask"]), device="cuda")
    c_tot0 = model.c[1:-1, 1:-1].double().sum().item()
    print(f"setup {t_setup:.1f} s; radii {model.radii}", flush=True)
    info = model.step(10)
    assert info["converged"]
    u0 = max(model.u.abs().max().item(), model.v.abs().max().item())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    n_steps = 190
    info = model.step(n_steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t1
    print(f"{n_steps} steps {dt:.2f} s = {n_steps / dt:.1f} steps/s; "
          f"{3 * n_steps * 3 / dt:.0f} modal solves/s; budgets {model.budgets}; "
          f"iterations/solve {info['iterations'] / max(info['solves'], 1):.2f}; "
          f"retries {model.retries}", flush=True)
    assert info["converged"]
    q = model.q[:, 1:-1, 1:-1]
    assert torch.isfinite(q).all()
    land = ~wet
    assert (q[:, land] == 0).all() and (model.psi[:, 1:-1, 1:-1][:, land] == 0).all()
    c = model.c[1:-1, 1:-1]
    assert torch.isfinite(c).all() and (c[land] == 0).all()
    c_tot = c.double().sum().item()
    # flux form: the total changes only by FP32 rounding of the stage sums
    assert abs(c_tot - c_tot0) <= 1e-4 * c.double().abs().sum().item(), (c_tot0, c_tot)
    # speeds stay below the CFL limit of the run
    umax = max(model.u.abs().max().item(), model.v.abs().max().item())
    print(f"max tracer speed {u0:.2f} m/s after 10 steps, {umax:.2f} m/s at the end", flush=True)
    assert umax * model.dt / model.dx < 1.0, umax
