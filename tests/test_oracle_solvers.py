"""Pin the oracle's elliptic pieces: operators as sparse matrices, transfers,
PCG against direct sparse solves, QG step pieces (CPU only)."""

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import fvx


def _mask(ny, nx, seed):
    rng = np.random.default_rng(seed)
    m = (rng.random((ny, nx)) > 0.2).astype(float)
    return m


@pytest.mark.parametrize("kind", ["noflux", "dirichlet"])
def test_operator_symmetric_negative_definite(kind):
    ny, nx = 12, 10
    m = _mask(ny, nx, 0)
    coeff = np.random.default_rng(1).uniform(0.5, 2, (ny, nx)) if kind == "noflux" else None
    lev = fvx.build_multigrid(m, 0.3, 0.2, 2.0, coeff, n_levels=1, kind=kind)[0]
    A = fvx.operator_matrix(lambda u: fvx.apply_operator(u, lev), (ny, nx)).toarray()
    wet = m.ravel() > 0
    Aw = A[np.ix_(wet, wet)]
    np.testing.assert_allclose(Aw, Aw.T, atol=1e-12)
    assert np.linalg.eigvalsh(Aw).max() < 0


def test_dirichlet_equals_masked_laplacian_with_dry_ring():
    ny, nx = 14, 11
    m = np.zeros((ny + 2, nx + 2))
    m[1:-1, 1:-1] = _mask(ny, nx, 3)
    psi = np.random.default_rng(4).standard_normal(m.shape)
    ref = fvx.masked_laplacian(psi, m, 0.5, 0.25, 3.0)
    lev = fvx.build_multigrid(m[1:-1, 1:-1], 0.5, 0.25, 3.0, n_levels=1, kind="dirichlet")[0]
    np.testing.assert_allclose(fvx.apply_operator(psi[1:-1, 1:-1], lev), ref[1:-1, 1:-1], atol=1e-12)
    A, _ = fvx.dirichlet_helmholtz_matrix(m[1:-1, 1:-1], 0.5, 0.25, 3.0)
    wet = m[1:-1, 1:-1] > 0
    u = psi[1:-1, 1:-1] * wet
    np.testing.assert_allclose(A @ u[wet], ref[1:-1, 1:-1][wet], atol=1e-12)


def test_transfers_constants_and_masks():
    """Reference tests/test_multigrid.py:137-196."""
    mf = np.ones((16, 16))
    mc = np.ones((8, 8))
    np.testing.assert_allclose(fvx.restrict(np.full((16, 16), 3.14), mf, mc), 3.14, atol=1e-12)
    np.testing.assert_allclose(fvx.prolongate(np.full((8, 8), 2.71), mc, mf), 2.71, atol=1e-12)
    mf2 = mf.copy(); mf2[:, :8] = 0
    mc2 = mc.copy(); mc2[:, :4] = 0
    out = fvx.restrict(np.full((16, 16), 5.0), mf2, mc2)
    assert np.all(out[mc2 == 0] == 0)


@pytest.mark.parametrize("kind,flexible", [("noflux", True), ("dirichlet", True),
                                           ("dirichlet", False)])
def test_pcg_matches_direct(kind, flexible):
    ny, nx = 32, 24
    m = _mask(ny, nx, 5)
    m[:2] = 1  # keep a wet strip connected to the edge
    coeff = np.random.default_rng(6).uniform(0.5, 2, (ny, nx)) if kind == "noflux" else None
    levels = fvx.build_multigrid(m, 1 / nx, 1 / ny, 4.0, coeff, kind=kind)
    b = np.random.default_rng(7).standard_normal((ny, nx)) * m
    A = fvx.operator_matrix(lambda u: fvx.apply_operator(u, levels[0]), (ny, nx))
    wet = m.ravel() > 0
    xd = np.zeros(ny * nx)
    xd[wet] = spla.spsolve(A[wet][:, wet].tocsc(), b.ravel()[wet])
    x, it, res, ok = fvx.pcg(lambda u: fvx.apply_operator(u, levels[0]), b,
                             precond=lambda r: fvx.v_cycle(levels, np.zeros_like(r), r, 0, 2, 2, 20),
                             rtol=1e-12, atol=0, max_steps=200, flexible=flexible)
    assert ok and res <= 2e-12 * np.linalg.norm(b)
    np.testing.assert_allclose(x.ravel(), xd, atol=1e-9 * np.abs(xd).max())


def test_planted_pcg_sign_error_is_caught():
    """A sign slip in the residual update diverges from the direct solve."""
    ny = nx = 16
    lev = fvx.build_multigrid(np.ones((ny, nx)), 1 / nx, 1 / ny, 4.0, kind="dirichlet")
    b = np.random.default_rng(1).standard_normal((ny, nx))
    A = lambda u: fvx.apply_operator(u, lev[0])
    x, it, res, ok = fvx.pcg(lambda u: -A(u), b, rtol=1e-10, atol=0, max_steps=50)
    assert not (ok and res <= 1e-9 * np.linalg.norm(b)) or np.linalg.norm(A(x) - b) > 1e-6


def test_qg_mg_solver_matches_direct():
    n = 48
    p = fvx.QGParams(n, n, 5.12e6, 5.12e6)
    F = p.forcing()
    wet = np.ones((n, n), bool)
    q0 = np.zeros((n + 2, n + 2))
    q0[1:-1, 1:-1] = np.random.default_rng(2).standard_normal((n, n)) * 1e-5
    levels = fvx.build_multigrid(wet.astype(float), p.dx, p.dy, p.lambda_, kind="dirichlet")

    def solver(rhs):
        x, it, res, ok = fvx.pcg(lambda u: fvx.apply_operator(u, levels[0]), rhs,
                                 precond=lambda r: fvx.v_cycle(levels, np.zeros_like(r), r, 0, 2, 2, 20, 0.8),
                                 rtol=1e-13, atol=0, max_steps=100)
        assert ok
        return x

    a = fvx.qg_step(q0, p, F, wet, solver)
    b = fvx.qg_step(q0, p, F, wet)
    np.testing.assert_allclose(a, b, atol=1e-10 * np.abs(b).max())
    assert np.all(a[0] == 0) and np.all(a[:, -1] == 0)


def test_flexible_beta_needed_for_nonsymmetric_vcycle():
    """The mask-weighted restriction is not the prolongation's transpose
    (multigrid.py:464-603), so the V-cycle preconditioner is non-symmetric on
    masked variable-coefficient problems: Fletcher-Reeves CG stalls where the
    flexible (Polak-Ribiere) beta converges."""
    ny, nx = 32, 24
    m = _mask(ny, nx, 5)
    m[:2] = 1
    coeff = np.random.default_rng(6).uniform(0.5, 2, (ny, nx))
    levels = fvx.build_multigrid(m, 1 / nx, 1 / ny, 4.0, coeff, kind="noflux")
    b = np.random.default_rng(7).standard_normal((ny, nx)) * m
    A = lambda u: fvx.apply_operator(u, levels[0])
    M = lambda r: fvx.v_cycle(levels, np.zeros_like(r), r, 0, 2, 2, 20)
    _, it_f, _, ok_f = fvx.pcg(A, b, precond=M, rtol=1e-12, atol=0, max_steps=200, flexible=True)
    _, it_p, _, ok_p = fvx.pcg(A, b, precond=M, rtol=1e-12, atol=0, max_steps=200, flexible=False)
    assert ok_f and (not ok_p or it_p > it_f)


@pytest.mark.parametrize("kind", ["noflux", "dirichlet"])
def test_level_matrix_matches_brute_force(kind):
    ny, nx = 9, 12
    m = _mask(ny, nx, 11)
    coeff = np.random.default_rng(12).uniform(0.5, 2, (ny, nx)) if kind == "noflux" else None
    lev = fvx.build_multigrid(m, 0.3, 0.7, 1.5, coeff, n_levels=1, kind=kind)[0]
    A1 = fvx.operator_matrix(lambda u: fvx.apply_operator(u, lev), (ny, nx)).toarray()
    A2 = fvx.level_matrix(lev).toarray()
    np.testing.assert_allclose(A2, A1, atol=1e-12 * np.abs(A1).max())
