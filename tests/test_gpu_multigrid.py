"""GPU parity: multigrid hierarchy, transfers, smoother, V-cycle and PCG.

FP64 single operators match the oracle to 1e-12 relative (max-norm scaled by
max|oracle|); composed V-cycles to 1e-10; FP32 to 1e-5 (operators) and 1e-4
(V-cycles).  Converged solves are checked against direct sparse solves and
by their FP64 true residual.
"""

import numpy as np
import pytest
import torch

import finitevolx_amd as fx
from finitevolx_amd import solvers as S
from finitevolx_amd.inputs import basin_mask
from oracle import fvx
from tests.gpu_util import assert_close_scaled, need_gpu, random_mask, to_np

pytestmark = pytest.mark.gpu


def _setup(ny, nx, kind, masked, varcoef, seed=0, lam=10.0, basin=False):
    """Ragged random masks exercise every operator branch; solves use a basin
    (one connected coastline): on ragged masks with many isolated pockets the
    V-cycle preconditioner itself stalls (the oracle's PCG too)."""
    rng = np.random.default_rng(seed)
    if masked and basin:
        m = basin_mask(ny, nx, seed).astype(np.float64)
        m[0, :] = m[-1, :] = m[:, 0] = m[:, -1] = 0.0
    elif masked:
        m = random_mask(ny, nx, seed + 1, 0.2).astype(np.float64)
    else:
        m = np.ones((ny, nx))
    if masked:  # keep the array edge partly wet so edge handling is exercised
        m[0, ny // 3:] = 1.0
    coeff = rng.uniform(0.5, 2.0, (ny, nx)) if varcoef else None
    dx, dy = 1.0 / nx, 1.3 / ny
    return m, coeff, dx, dy, lam


CASES = [
    (16, 16, "noflux", False, False), (32, 32, "noflux", True, True), (16, 32, "dirichlet", False, False),
    (128, 96, "noflux", True, True), (256, 192, "dirichlet", True, False), (192, 256, "noflux", False, True),
]


@pytest.mark.parametrize("ny,nx,kind,masked,varcoef", CASES)
def test_operator_transfers_jacobi(ny, nx, kind, masked, varcoef):
    need_gpu()
    m, coeff, dx, dy, lam = _setup(ny, nx, kind, masked, varcoef)
    levels = fvx.build_multigrid(m, dx, dy, lam, coeff, kind=kind)
    for prec, tol in (("float64", 1e-12), ("float32", 1e-5)):
        mg = S.build_multigrid_solver(m, dx, dy, lam, coeff, kind=kind, precision=prec,
                                      n_pre=2, n_post=2)
        assert mg.n_levels == len(levels)
        rng = np.random.default_rng(3)
        for li, lev in enumerate(levels):
            u = rng.standard_normal(lev.mask.shape)
            assert_close_scaled(to_np(S.apply_operator(u, mg.levels[li])),
                                fvx.apply_operator(u, lev), tol, f"A level {li} {prec}")
            b = rng.standard_normal(lev.mask.shape)
            for n in (1, 2, 3):
                got = S.weighted_jacobi(u, b, mg.levels[li], 0.95, n)
                assert_close_scaled(to_np(got), fvx.weighted_jacobi(u, b, lev, 0.95, n), 10 * tol,
                                    f"jacobi{n} level {li} {prec}")
            if li + 1 < len(levels):
                cl = levels[li + 1]
                R = fvx.restrict_dirichlet if lev.kind == "dirichlet" else fvx.restrict
                assert_close_scaled(to_np(S.restrict(u, mg.levels[li])),
                                    R(u, lev.mask, cl.mask), tol, f"restrict {li}")
                vc = rng.standard_normal(cl.mask.shape)
                P = fvx.prolongate_dirichlet if lev.kind == "dirichlet" else fvx.prolongate
                assert_close_scaled(to_np(S.prolongate(vc, mg.levels[li])),
                                    P(vc, cl.mask, lev.mask), tol, f"prolong {li}")


@pytest.mark.parametrize("ny,nx,kind,masked,varcoef", CASES)
@pytest.mark.parametrize("npre,npost", [(6, 6), (2, 2), (1, 3)])
def test_v_cycle_matches_oracle(ny, nx, kind, masked, varcoef, npre, npost):
    need_gpu()
    m, coeff, dx, dy, lam = _setup(ny, nx, kind, masked, varcoef, seed=4)
    levels = fvx.build_multigrid(m, dx, dy, lam, coeff, kind=kind)
    rng = np.random.default_rng(5)
    b = rng.standard_normal((ny, nx)) * m
    u0 = rng.standard_normal((ny, nx))
    kw = dict(n_pre=npre, n_post=npost, n_coarse=30, omega=0.9)
    mg = S.build_multigrid_solver(m, dx, dy, lam, coeff, kind=kind, **kw)
    ref0 = fvx.v_cycle(levels, np.zeros_like(b), b, **kw)
    assert_close_scaled(to_np(mg.v_cycle(None, b)), ref0, 1e-10, "V(0,b) f64")
    ref1 = fvx.v_cycle(levels, u0, b, **kw)
    assert_close_scaled(to_np(mg.v_cycle(u0, b)), ref1, 1e-10, "V(u0,b) f64")
    mg32 = S.build_multigrid_solver(m, dx, dy, lam, coeff, kind=kind, precision="float32", **kw)
    got32 = mg32.v_cycle(None, torch.tensor(b, dtype=torch.float32))
    assert_close_scaled(to_np(got32), ref0, 1e-4, "V(0,b) f32")


@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("ny,nx,nl", [(512, 512, 3), (384, 640, 2)])
def test_wall_corrected_fp32_v_cycle_matches_oracle(ny, nx, nl, masked):
    """The coarse-level wall correction (every level tiled, wave-tile
    kernels, ghost value g_l u beside land and the edge) against the oracle's
    corrected hierarchy, from zero and from a nonzero initial guess; it is
    refused where a coarse level would fall to the single-workgroup kernel,
    and the corrected levels' operators are not exposed."""
    need_gpu()
    m = basin_mask(ny, nx, 31).astype(np.float64) if masked else np.ones((ny, nx))
    dx, dy, lam = 1.0 / nx, 1.0 / ny, 40.0
    kw = dict(n_pre=3, n_post=3, n_coarse=3, omega=0.85)
    levels = fvx.build_multigrid(m, dx, dy, lam, n_levels=nl, kind="dirichlet", wall=True)
    mg = S.build_multigrid_solver(m, dx, dy, lam, n_levels=nl, kind="dirichlet",
                                  precision="float32", wall=True, **kw)
    rng = np.random.default_rng(22)
    b = rng.standard_normal((ny, nx)) * m
    u0 = rng.standard_normal((ny, nx)) * m
    for u in (None, u0):
        ref = fvx.v_cycle(levels, np.zeros_like(b) if u is None else u, b, **kw)
        got = mg.v_cycle(None if u is None else torch.tensor(u, dtype=torch.float32),
                         torch.tensor(b, dtype=torch.float32))
        assert_close_scaled(to_np(got), ref, 1e-4, f"wall V-cycle u0={u is not None}")
    # without the correction the cycle differs (the test sees the correction)
    plain = fvx.v_cycle(fvx.build_multigrid(m, dx, dy, lam, n_levels=nl, kind="dirichlet"),
                        np.zeros_like(b), b, **kw)
    corr = fvx.v_cycle(levels, np.zeros_like(b), b, **kw)
    assert np.abs(plain - corr).max() > 1e-3 * np.abs(plain).max()
    with pytest.raises(ValueError):
        S.apply_operator(torch.zeros(mg.levels[1].shape, dtype=torch.float32), mg.levels[1])
    deep = S.build_multigrid_solver(m, dx, dy, lam, kind="dirichlet", precision="float32", **kw)
    with pytest.raises(ValueError):
        from finitevolx_amd import _lib
        _lib.call("fvx_mg_plan_set_wall", deep._h, 1)


@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("ny,nx,nl,sweeps", [(512, 512, 3, (3, 3, 3)), (384, 640, 2, (2, 1, 3)),
                                             (256, 320, 5, (2, 3, 5))])
def test_chebyshev_fp32_v_cycle_matches_oracle(ny, nx, nl, sweeps, masked):
    """Chebyshev weights on [0.1, 2] for the sweeps of the V-cycle
    (pre-smoothing and bottom ascending, post-smoothing descending) against
    the oracle's V-cycle with the same weights, from zero and from a nonzero
    initial guess: every level on the wave-tile kernels, or (256 x 320, 5
    levels) the last two in the single-workgroup coarse kernel with 5 bottom
    sweeps; the FP64 plan (LDS tile kernels) to 1e-10; lo = 0 restores the
    uniform weight; more than 3 sweeps a leg is refused."""
    need_gpu()
    from finitevolx_amd import _lib

    m = basin_mask(ny, nx, 31).astype(np.float64) if masked else np.ones((ny, nx))
    dx, dy, lam = 1.0 / nx, 1.0 / ny, 40.0
    kw = dict(n_pre=sweeps[0], n_post=sweeps[1], n_coarse=sweeps[2], omega=0.85)
    levels = fvx.build_multigrid(m, dx, dy, lam, n_levels=nl, kind="dirichlet")
    mg = S.build_multigrid_solver(m, dx, dy, lam, n_levels=nl, kind="dirichlet",
                                  precision="float32", chebyshev=0.1, **kw)
    rng = np.random.default_rng(23)
    b = rng.standard_normal((ny, nx)) * m
    u0 = rng.standard_normal((ny, nx)) * m
    for u in (None, u0):
        ref = fvx.v_cycle(levels, np.zeros_like(b) if u is None else u, b, cheb=0.1, **kw)
        got = mg.v_cycle(None if u is None else torch.tensor(u, dtype=torch.float32),
                         torch.tensor(b, dtype=torch.float32))
        assert_close_scaled(to_np(got), ref, 1e-4, f"Chebyshev V-cycle u0={u is not None}")
    plain = fvx.v_cycle(levels, np.zeros_like(b), b, **kw)
    assert np.abs(plain - fvx.v_cycle(levels, np.zeros_like(b), b, cheb=0.1, **kw)).max() > \
        1e-3 * np.abs(plain).max()
    _lib.call("fvx_mg_plan_set_chebyshev", mg._h, 0.0)
    assert_close_scaled(to_np(mg.v_cycle(None, torch.tensor(b, dtype=torch.float32))), plain,
                        1e-4, "uniform weights restored")
    mg64 = S.build_multigrid_solver(m, dx, dy, lam, n_levels=nl, kind="dirichlet",
                                    precision="float64", chebyshev=0.1, **kw)
    ref0 = fvx.v_cycle(levels, np.zeros_like(b), b, cheb=0.1, **kw)
    assert_close_scaled(to_np(mg64.v_cycle(None, b)), ref0, 1e-10, "FP64 Chebyshev V-cycle")
    with pytest.raises(ValueError):
        S.build_multigrid_solver(m, dx, dy, lam, n_levels=nl, kind="dirichlet",
                                 precision="float32", chebyshev=0.1, n_pre=6, n_post=6,
                                 n_coarse=3)


@pytest.mark.parametrize("ny,nx,masked", [(128, 96, True), (192, 256, False)])
def test_chebyshev_noflux_v_cycle_matches_oracle(ny, nx, masked):
    """Chebyshev weights on the no-flux variable-coefficient operator (LDS
    tile kernels and the coarse kernel with 6 bottom sweeps) against the
    oracle: FP64 to 1e-10, FP32 to 1e-4."""
    need_gpu()
    m, coeff, dx, dy, lam = _setup(ny, nx, "noflux", masked, True, seed=6)
    levels = fvx.build_multigrid(m, dx, dy, lam, coeff, kind="noflux")
    rng = np.random.default_rng(7)
    b = rng.standard_normal((ny, nx)) * m
    u0 = rng.standard_normal((ny, nx))
    kw = dict(n_pre=2, n_post=2, n_coarse=6, omega=0.8)
    for prec, tol in (("float64", 1e-10), ("float32", 1e-4)):
        mg = S.build_multigrid_solver(m, dx, dy, lam, coeff, kind="noflux", precision=prec,
                                      chebyshev=0.2, **kw)
        dt = torch.float64 if prec == "float64" else torch.float32
        for u in (None, u0):
            ref = fvx.v_cycle(levels, np.zeros_like(b) if u is None else u, b, cheb=0.2, **kw)
            got = mg.v_cycle(None if u is None else torch.tensor(u, dtype=dt),
                             torch.tensor(b, dtype=dt))
            assert_close_scaled(to_np(got), ref, tol, f"noflux Chebyshev {prec} u0={u is not None}")


@pytest.mark.parametrize("ny,nx", [(256, 192), (200, 168), (96, 520)])
@pytest.mark.parametrize("npre,npost", [(2, 2), (1, 3), (6, 6)])
def test_all_wet_dirichlet_fp32_v_cycle_matches_oracle(ny, nx, npre, npost):
    """FP32 levels without a land mask under the Dirichlet operator run the
    wavefront-tile kernels (registers + lane moves); sizes whose tile counts
    are ragged and hierarchies deep enough to tile several levels, from a zero
    and from a nonzero initial guess."""
    need_gpu()
    m = np.ones((ny, nx))
    dx, dy, lam = 1.0 / nx, 1.3 / ny, 10.0
    levels = fvx.build_multigrid(m, dx, dy, lam, kind="dirichlet")
    rng = np.random.default_rng(21)
    b = rng.standard_normal((ny, nx))
    u0 = rng.standard_normal((ny, nx))
    kw = dict(n_pre=npre, n_post=npost, n_coarse=30, omega=0.9)
    mg32 = S.build_multigrid_solver(m, dx, dy, lam, kind="dirichlet", precision="float32", **kw)
    got0 = mg32.v_cycle(None, torch.tensor(b, dtype=torch.float32))
    assert_close_scaled(to_np(got0), fvx.v_cycle(levels, np.zeros_like(b), b, **kw), 1e-5,
                        "V(0,b) f32")
    got1 = mg32.v_cycle(torch.tensor(u0, dtype=torch.float32), torch.tensor(b, dtype=torch.float32))
    assert_close_scaled(to_np(got1), fvx.v_cycle(levels, u0, b, **kw), 1e-5, "V(u0,b) f32")


@pytest.mark.parametrize("ny,nx", [(256, 192), (200, 168), (96, 520)])
@pytest.mark.parametrize("nsw", [2, 3])
def test_all_wet_dirichlet_fp32_pcg_iterates_match_oracle(ny, nx, nsw):
    """The first flexible-PCG iterates x_1..x_4 of the FP32 solver (V-cycle
    whose last kernel also forms A z and the dot products, and whose first
    kernel applies the previous step; 2 or 3 fused sweeps; widths that are
    not multiples of the 54- and 56-column tile interiors) against the
    oracle's PCG with the same V-cycle: any error in z, A z, the fused step or
    the step scalars shows at the 1e-2 level, FP32 rounding at 1e-6."""
    need_gpu()
    m = np.ones((ny, nx))
    dx, dy, lam = 1.0 / nx, 1.3 / ny, 10.0
    levels = fvx.build_multigrid(m, dx, dy, lam, kind="dirichlet")
    b = np.random.default_rng(22).standard_normal((ny, nx))
    kw = dict(n_pre=nsw, n_post=nsw, n_coarse=4, omega=0.85)
    M = lambda r: fvx.v_cycle(levels, np.zeros_like(r), r, **kw)
    A = lambda u: fvx.apply_operator(u, levels[0])
    mg32 = S.build_multigrid_solver(m, dx, dy, lam, kind="dirichlet", precision="float32", **kw)
    for k in (1, 2, 3, 4):
        xr, _, _, _ = fvx.pcg(A, b, precond=M, rtol=1e-12, atol=0.0, max_steps=k)
        x, info = mg32.pcg(torch.tensor(b, dtype=torch.float32), rtol=1e-12, atol=0.0,
                           max_steps=k)
        assert info.iterations == k
        assert_close_scaled(to_np(x).astype(np.float64), xr, 1e-4, f"x_{k}")


def test_solver_call_and_reference_thresholds():
    """Mirrors reference tests/test_multigrid.py:222-301 thresholds."""
    need_gpu()
    ny = nx = 32
    rng = np.random.default_rng(11)
    rhs = rng.standard_normal((ny, nx))
    mg = S.build_multigrid_solver(np.ones((ny, nx)), 1 / nx, 1 / ny, lambda_=10.0, n_cycles=20)
    u = mg(rhs)
    lev = fvx.build_multigrid(np.ones((ny, nx)), 1 / nx, 1 / ny, 10.0)
    ref = fvx.mg_solve(lev, rhs, n_cycles=20)
    assert_close_scaled(to_np(u), ref, 1e-9, "20 cycles")
    res = np.linalg.norm(rhs - fvx.apply_operator(to_np(u), lev[0])) / np.linalg.norm(rhs)
    assert res < 1e-2
    with pytest.raises(ValueError, match="not divisible"):
        S.build_multigrid_solver(np.ones((15, 15)), 1.0, 1.0, n_levels=2)


def test_masked_outputs_zero_outside():
    need_gpu()
    ny = nx = 16
    jj, ii = np.meshgrid(np.linspace(-.5, .5, ny), np.linspace(-.5, .5, nx), indexing="ij")
    m = (jj**2 + ii**2 < 0.16).astype(np.float64)
    mg = S.build_multigrid_solver(m, 1 / nx, 1 / ny, n_levels=1, n_cycles=10)
    rhs = np.random.default_rng(7).standard_normal((ny, nx)) * m
    u = to_np(mg(rhs))
    assert np.all(u[m == 0] == 0.0)
    Au = to_np(S.apply_operator(np.ones((ny, nx)), mg.levels[0]))
    assert np.all(Au[m == 0] == 0.0)


@pytest.mark.parametrize("ny,nx,kind,masked,varcoef", CASES)
def test_pcg_converges_to_direct_solution(ny, nx, kind, masked, varcoef):
    need_gpu()
    m, coeff, dx, dy, lam = _setup(ny, nx, kind, masked, varcoef, seed=8, basin=True)
    rng = np.random.default_rng(9)
    b = rng.standard_normal((ny, nx)) * m
    lev = fvx.build_multigrid(m, dx, dy, lam, coeff, kind=kind)
    A = fvx.level_matrix(lev[0])
    wet = m.ravel() > 0
    import scipy.sparse.linalg as spla
    xd = np.zeros(ny * nx)
    xd[wet] = spla.spsolve(A[wet][:, wet].tocsc(), b.ravel()[wet])
    xd = xd.reshape(ny, nx)
    mg = S.build_multigrid_solver(m, dx, dy, lam, coeff, kind=kind, n_pre=2, n_post=2)
    x, info = fx.solve_cg(S.MultigridOperator(mg), b, preconditioner=S.make_multigrid_preconditioner(mg),
                          rtol=1e-11, atol=0.0, max_steps=200)
    assert info.converged, info
    assert_close_scaled(to_np(x), xd, 1e-8, "pcg vs direct")
    # the reported residual is the true residual
    r = np.linalg.norm(b - fvx.apply_operator(to_np(x), lev[0]))
    assert r <= 1.01e-11 * np.linalg.norm(b) * 1.5


def test_solve_cg_with_callables_polls_every_few_iterations():
    """solve_cg on plain callables (torch vector algebra): the host reads
    the device-side stopping test every 8 iterations, with the same iterate
    and iteration count as a test after every iteration, and the solution
    of the masked Dirichlet Helmholtz problem."""
    need_gpu()
    ny, nx = 40, 56
    m = random_mask(ny, nx, 6).astype(np.float64)
    q = np.random.default_rng(7).standard_normal((ny, nx)) * m
    mt = torch.tensor(m, device="cuda")
    A = lambda x: S.masked_laplacian(x, mt, 0.1, 0.2, 3.0)  # noqa: E731
    Mj = lambda r: r / (-2 / 0.01 - 2 / 0.04 - 3.0)  # noqa: E731
    x8, i8 = fx.solve_cg(A, q, preconditioner=Mj, rtol=1e-10, atol=0.0, max_steps=400)
    x1, i1 = S._generic_pcg(A, q, None, Mj, 1e-10, 0.0, 400, check_every=1)
    assert i8.converged and i1.converged
    assert i8.iterations == i1.iterations, (i8, i1)
    assert torch.equal(x8, x1)
    assert_close_scaled(to_np(x8), fvx.dirichlet_helmholtz_direct(q, m, 0.1, 0.2, 3.0), 1e-8, "cg")
    xs, is_ = fx.solve_cg(A, q, preconditioner=Mj, rtol=1e-10, atol=0.0, max_steps=5)
    assert not is_.converged and is_.iterations == 5


@pytest.mark.parametrize("kind", ["dirichlet", "noflux"])
def test_mixed_precision_certifies_fp64_residual(kind):
    need_gpu()
    ny, nx = 256, 192
    m, coeff, dx, dy, lam = _setup(ny, nx, kind, True, kind == "noflux", seed=12, basin=True)
    rng = np.random.default_rng(13)
    b = rng.standard_normal((ny, nx)) * m
    lev = fvx.build_multigrid(m, dx, dy, lam, coeff, kind=kind)
    mix = S.build_multigrid_solver(m, dx, dy, lam, coeff, kind=kind, precision="mixed",
                                   n_pre=2, n_post=2)
    x, info = mix.pcg(b, rtol=1e-10, atol=0.0, max_steps=60)
    assert info.converged
    r = np.linalg.norm(b - fvx.apply_operator(to_np(x), lev[0]))
    assert r <= 1e-10 * np.linalg.norm(b) * 1.0001, (r, info)
    assert mix.last_info.outer >= 2  # FP32 inner solves alone cannot reach 1e-10
    f64 = S.build_multigrid_solver(m, dx, dy, lam, coeff, kind=kind, n_pre=2, n_post=2)
    x64, i64 = f64.pcg(b, rtol=1e-10, atol=0.0, max_steps=60)
    assert i64.converged
    assert_close_scaled(to_np(x), to_np(x64), 1e-8, "mixed vs fp64")


def test_masked_laplacian_and_pv_inversion():
    need_gpu()
    ny, nx = 34, 50
    rng = np.random.default_rng(2)
    psi = rng.standard_normal((ny, nx))
    m = random_mask(ny, nx, 4).astype(np.float64)
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        got = S.masked_laplacian(torch.tensor(psi, dtype=dt), torch.tensor(m, dtype=dt), 0.1, 0.2, 3.0)
        assert_close_scaled(to_np(got), fvx.masked_laplacian(psi, m, 0.1, 0.2, 3.0), tol, "mlap")
    q = rng.standard_normal((ny, nx)) * m
    psi_cg = to_np(S.pv_inversion(q, 0.1, 0.2, 3.0, method="cg", mask=m))
    # exact solve of the masked (Dirichlet) Helmholtz problem
    ref = fvx.dirichlet_helmholtz_direct(q, m, 0.1, 0.2, 3.0)
    assert_close_scaled(psi_cg, ref, 1e-8, "pv_inversion cg")
    lam = np.array([1.0, 5.0])
    q2 = np.stack([q, 2 * q])
    out = to_np(S.pv_inversion(q2, 0.1, 0.2, lam, method="cg", mask=m))
    assert_close_scaled(out[1], fvx.dirichlet_helmholtz_direct(2 * q, m, 0.1, 0.2, 5.0), 1e-8, "layer 1")
    with pytest.raises(ValueError):
        S.pv_inversion(q2, 0.1, 0.2, np.array([1.0, 2.0, 3.0]), method="cg", mask=m)


def test_large_bottom_level_uses_coarse_kernel():
    """130x130 gives levels 130^2 and 65^2 (> the coarse-kernel size)."""
    need_gpu()
    n = 130
    m = np.ones((n, n))
    lev = fvx.build_multigrid(m, 1 / n, 1 / n, 5.0)
    assert len(lev) == 2
    mg = S.build_multigrid_solver(m, 1 / n, 1 / n, 5.0, n_pre=2, n_post=2, n_coarse=10)
    assert mg.coarse_level == 1
    b = np.random.default_rng(0).standard_normal((n, n))
    ref = fvx.v_cycle(lev, np.zeros_like(b), b, n_pre=2, n_post=2, n_coarse=10)
    assert_close_scaled(to_np(mg.v_cycle(None, b)), ref, 1e-10, "V 130")


def test_kernel_bench_down_step_mode():
    """fvx_mg_kernel_bench mode 3 (the level-0 down kernel with the previous
    PCG step) runs on FP32 Dirichlet plans, all-wet or with a land mask, and
    is refused, with the library's invalid-argument error, on a no-flux
    plan."""
    need_gpu()
    from finitevolx_amd import _lib

    _lib.extra_signatures({"fvx_mg_kernel_bench": [_lib._vp, _lib._i, _lib._i, _lib._i, _lib._vp]})
    ny = nx = 256
    kw = dict(n_pre=3, n_post=3, n_coarse=4, omega=0.85)
    wet = S.build_multigrid_solver(np.ones((ny, nx)), 1.0 / nx, 1.0 / ny, 10.0, kind="dirichlet",
                                   precision="mixed", **kw)
    for reps in (0, 3):
        _lib.call("fvx_mg_kernel_bench", wet._h, 3, 32, reps, _lib.stream())
    torch.cuda.synchronize()
    m = np.ones((ny, nx))
    m[:, :9] = 0.0
    masked = S.build_multigrid_solver(m, 1.0 / nx, 1.0 / ny, 10.0, kind="dirichlet",
                                      precision="mixed", **kw)
    for reps in (0, 3):
        _lib.call("fvx_mg_kernel_bench", masked._h, 3, 32, reps, _lib.stream())
    torch.cuda.synchronize()
    noflux = S.build_multigrid_solver(m, 1.0 / nx, 1.0 / ny, 10.0, kind="noflux",
                                      precision="mixed", **kw)
    _lib.call("fvx_mg_kernel_bench", noflux._h, 3, 32, 0, _lib.stream())
    with pytest.raises(ValueError):
        _lib.call("fvx_mg_kernel_bench", noflux._h, 3, 32, 1, _lib.stream())


@pytest.mark.parametrize("ny,nx,kind,masked,varcoef", [CASES[3], CASES[4], CASES[5]])
def test_v_cycle_from_coarser_levels(ny, nx, kind, masked, varcoef):
    """MultigridSolver.v_cycle(u, rhs, level_idx) from every level below the
    finest (multigrid.py:808-883) against the oracle's recursion from that
    level: FP64 1e-10, FP32 1e-4."""
    need_gpu()
    m, coeff, dx, dy, lam = _setup(ny, nx, kind, masked, varcoef, seed=6)
    kw = dict(n_pre=2, n_post=2, n_coarse=30, omega=0.9)
    levels = fvx.build_multigrid(m, dx, dy, lam, coeff, kind=kind)
    mg = S.build_multigrid_solver(m, dx, dy, lam, coeff, kind=kind, **kw)
    mg32 = S.build_multigrid_solver(m, dx, dy, lam, coeff, kind=kind, precision="float32", **kw)
    rng = np.random.default_rng(7)
    for li in range(1, len(levels)):
        shape = levels[li].mask.shape
        b = rng.standard_normal(shape) * levels[li].mask
        u0 = rng.standard_normal(shape)
        ref = fvx.v_cycle(levels, u0, b, li, **kw)
        assert_close_scaled(to_np(mg.v_cycle(u0, b, level_idx=li)), ref, 1e-10, f"level {li}")
        ref0 = fvx.v_cycle(levels, np.zeros(shape), b, li, **kw)
        got = mg32.v_cycle(None, torch.tensor(b, dtype=torch.float32), level_idx=li)
        assert_close_scaled(to_np(got), ref0, 1e-4, f"level {li} f32")
    with pytest.raises(ValueError):
        mg.v_cycle(None, np.zeros((ny, nx)), level_idx=len(levels))
