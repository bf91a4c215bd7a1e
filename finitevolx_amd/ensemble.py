"""Ensembles of independent 1.5-layer QG basins of uneven sizes on one GPU.

Each member is the model of ``finitevolx_amd.qg`` (closed basin, its own grid
shape, land mask and deformation radius).  Every RK stage runs as one
work-list launch per operator over all members: the extrapolated warm start
(one elementwise launch over the members' contiguous streamfunction
buffers), the mixed-precision Helmholtz solve (``MultigridBatch``: each
operator of the solve one launch over every member), and the fused
tendency + stage combination (``fvx_qg_stage_batch_launch``).  A step is
replayed from a captured HIP graph; each stage's solve budget (refinement
steps and inner iterations, the largest over members) is calibrated on the
first steps, tightened to what the members use, and widened when a member
misses (the step is then redone from a copy of its input).  The reference
batches independent problems with ``vmap`` (``solvers/elliptic.py:532-541``,
``vertical/multilayer.py:99``); here the batch is the launch.
"""

from __future__ import annotations

import gc

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np
import torch

from finitevolx_amd import _lib
from finitevolx_amd.solvers import MultigridBatch, build_multigrid_solver
from finitevolx_amd.timestepping import lincomb

_vp, _i, _d = _lib._vp, _lib._i, _lib._d


class _StageDesc(C.Structure):
    _fields_ = [("qs", _vp), ("psi", _vp), ("qn", _vp), ("out", _vp), ("mask", _vp),
                ("Ny", _i), ("Nx", _i), ("dx", _d), ("dy", _d), ("beta", _d), ("Ly", _d),
                ("F0", _d), ("nu4", _d), ("a", _d), ("b", _d), ("cdt", _d)]


_lib.extra_signatures({
    "fvx_qg_stage_batch_create": [_i, C.POINTER(_StageDesc), _i, C.POINTER(_vp)],
    "fvx_qg_stage_batch_launch": [_vp, _vp],
    "fvx_qg_stage_batch_destroy": [_vp],
})

COEF = ((0.0, 1.0, 1.0), (0.75, 0.25, 0.25), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0))


@dataclass
class Member:
    """One basin: interior shape, grid spacing, deformation radius, optional
    interior wet mask and initial interior PV."""
    ny: int
    nx: int
    dx: float
    Ld: float
    q0: np.ndarray
    mask: np.ndarray | None = None


def hierarchy_depth(ny, nx, dx, Ld, max_coarse_cells=64 * 64):
    """Coarsen until the cell size reaches 2 Ld, within the halving rule of
    multigrid.py:50-76, and at least until the coarsest level has at most
    ``max_coarse_cells`` cells: a batched solve runs its coarse levels in one
    workgroup's LDS (the single-model QG plan stops at Ld / 2 and smooths a
    large bottom level on the tiled kernels instead)."""
    deepest, hy, hx = 1, ny, nx
    fits = 1 if ny * nx <= max_coarse_cells else None
    while hy % 2 == 0 and hx % 2 == 0 and hy // 2 >= 8 and hx // 2 >= 8:
        hy //= 2
        hx //= 2
        deepest += 1
        if fits is None and hy * hx <= max_coarse_cells:
            fits = deepest
    ratio = 2.0 * Ld / dx
    if not np.isfinite(ratio):  # an infinite radius (Poisson): the deepest hierarchy
        return deepest
    n = max(2, 1 + int(np.ceil(np.log2(max(ratio, 1.0)))), fits or deepest)
    return min(n, deepest)


def _budget(outer, inner, margin):
    """(refinement steps, inner iterations per step) from observed counts
    (reported for the first four steps; later steps repeat the fourth)."""
    ni = tuple(v + margin for v in inner)
    return outer, ni + (ni[-1],) * max(0, outer - len(ni))


class BudgetedSteps:
    """Graph-replayed rk3_ssp_step steps whose three stages each end in a
    batch of mixed-precision solves (``self.batch``, slot = stage) with a
    per-stage budget (refinement steps, inner iterations per step).

    The first three steps run eagerly; the third (warm) step's counts become
    the budgets.  Then each parity of the extrapolation buffers has a
    captured graph; after ``win_len`` converged steps each stage's budget
    comes down to the largest counts used in that window.  When a stage
    misses, the step is redone eagerly from the saved state with a generous
    budget for that stage, and its observed counts plus one iteration per
    refinement step become the new budgets (the window doubles, up to 16
    steps).  Subclasses define ``_enqueue_step(e, fresh)`` (which saves the
    state first), ``_restore()``, ``_build_lists()`` and ``_state()`` (the
    tensors a trial step may change)."""

    WIDE = (6, (48, 32, 16, 16, 16, 16))

    def _init_budgets(self):
        self.parity = 0
        self.fresh = True
        self.budgets = [(3, (12, 8, 8)) for _ in range(3)]
        self.graphs = {}
        self.steps_done = 0
        self.retries = 0
        self.recaptures = 0
        self.win = None
        self.win_steps = 0
        self.win_len = 4
        self._lists_ready = False
        self._recal = False

    def _status(self):
        """Per stage slot: (every member converged, refinement steps used,
        inner iterations used per step (largest over members), total inner
        iterations)."""
        out = []
        for k in range(3):
            no, _ = self.budgets[k]
            st = self.batch.status(slot=k, n_outer=no)
            ok = all(s[0].converged for s in st)
            outer = max(max(s[1] for s in st), 1)
            inner = tuple(max(max(s[2][o] for s in st), 1) for o in range(min(outer, 4)))
            out.append((ok, outer, inner, sum(s[0].iterations for s in st)))
        return out

    def _capture(self):
        # Unreachable plans freed by the garbage collector during a capture
        # would call hipFree, which ends the capture with an error: the
        # collector stays off while capturing.
        self.graphs = {}
        was_enabled = gc.isenabled()
        gc.disable()
        try:
            for e in (0, 1):
                g = torch.cuda.CUDAGraph()
                s = torch.cuda.Stream()
                with torch.cuda.stream(s):
                    with torch.cuda.graph(g, stream=s):
                        self._enqueue_step(e)
                self.graphs[e] = g
        finally:
            if was_enabled:
                gc.enable()
        self.recaptures += 1

    def step(self, nsteps: int = 1, graph: bool = True):
        """Advance by ``nsteps`` rk3_ssp_step steps.  Returns a dict of
        totals (solves, inner iterations, redone steps, converged)."""
        info = dict(solves=0, iterations=0, retries=0, converged=True)
        done = 0
        while done < nsteps:
            calib = self.steps_done < 3 or not graph or self._recal
            if calib:
                self._enqueue_step(self.parity, fresh=self.fresh)
            else:
                if not self.graphs:
                    if not self._lists_ready:
                        self._warm_lists()
                        self._lists_ready = True
                    self._capture()
                self.graphs[self.parity].replay()
            st = self._status()
            ok = all(s[0] for s in st)
            if not ok:
                # redo the step eagerly with a generous budget for the missed
                # stages; its counts then set the budgets
                for k, s in enumerate(st):
                    if not s[0]:
                        self.budgets[k] = self.WIDE
                if not calib:
                    self.win_len = min(16, 2 * self.win_len)
                self.win = None
                self.win_steps = 0
                self.graphs = {}
                self._recal = True
                self._restore()
                self.retries += 1
                info["retries"] += 1
                if self.retries > 50:
                    raise RuntimeError("batched solves keep missing their budgets")
                continue
            for s in st:
                info["solves"] += self.batch.n
                info["iterations"] += s[3]
            if self._recal:
                # observed counts with one iteration of margin per refinement step
                self.budgets = [_budget(s[1], s[2], 1) for s in st]
                self._recal = False
            elif calib and self.steps_done == 2:
                # budgets from the third (warm) step's counts
                self.budgets = [_budget(s[1], s[2], 0) for s in st]
            elif not calib:
                self._tighten(st)
            self.parity ^= 1
            self.fresh = False
            self.steps_done += 1
            done += 1
        return info

    def _warm_lists(self):
        # build every work-list and solver slot list eagerly, outside capture
        self._build_lists()
        saved = [t.clone() for t in self._state()]
        for e in (0, 1):
            self._enqueue_step(e)
        torch.cuda.synchronize()
        for t, sv in zip(self._state(), saved):
            t.copy_(sv)

    def _tighten(self, st):
        """After win_len converged graph steps, each stage's budgets come down
        to the largest counts its members used in that window."""
        used = [(s[1], s[2]) for s in st]
        if self.win is None:
            self.win = used
        else:
            merged = []
            for (ao, ai), (bo, bi) in zip(self.win, used):
                n = max(ao, bo)
                pad = lambda t: tuple(t) + (0,) * (n - len(t))  # noqa: E731
                merged.append((n, tuple(max(x, y) for x, y in zip(pad(ai), pad(bi)))))
            self.win = merged
        self.win_steps += 1
        if self.win_steps < self.win_len:
            return
        new = []
        for (no, ni), (wo, wi) in zip(self.budgets, self.win):
            o = min(no, max(1, wo))
            wi = tuple(wi) + (1,) * max(0, o - len(wi))
            new.append((o, tuple(min(ni[j], max(1, wi[j])) for j in range(o))))
        self.win = None
        self.win_steps = 0
        if new != self.budgets:
            self.budgets = new
            self.graphs = {}


def extrapolated_guess(ps, k, e, fresh):
    """Warm start of stage k's streamfunction over whole buffers (the
    extrapolation of qg.hip's stage_guess: psi of dt times the tendency,
    known exactly from the previous step's stages, taken linear in time).
    ``ps`` holds four buffers: the previous and next step-start solutions
    (which of ps[0] / ps[3] is which alternates with the parity ``e``) and
    the stage-1 and stage-2 solutions."""
    P, N = (ps[0], ps[3]) if e == 0 else (ps[3], ps[0])
    B1, B2 = ps[1], ps[2]
    if k == 0:
        if fresh:
            lincomb(1.0, B2, out=N)
        else:
            lincomb(2.0, B2, -1.0, P, out=N)
    elif k == 1:
        if fresh:
            lincomb(1.0, N, out=B1)
        else:
            lincomb(4.0, N, -2.0, B2, -1.0, B1, out=B1)
    else:
        if fresh:
            lincomb(1.0, B1, out=B2)
        else:
            lincomb(1.0, B1, -0.75, N, 0.25, P, out=B2, d=0.5, w=B2)


def stage_psi_buffer(ps, k, e):
    """The buffer stage k of parity e solves into."""
    N = ps[3] if e == 0 else ps[0]
    return (N, ps[1], ps[2])[k]


class QGEnsemble(BudgetedSteps):
    """Batched mixed-precision QG steps (FP64 PV state, FP32 tendencies and
    inner solves, FP64 refinement) of independent members.  ``chebyshev``:
    the lower end lo of [lo, 2] whose Chebyshev weights the members'
    V-cycle sweeps take (None: the uniform omega).  Measured on the
    96-member configuration: 86 ensemble steps/s uniform, 93 / 95 / 95 at lo
    = 0.2 / 0.25 / 0.3 (2.38 -> 2.09 inner iterations per solve at 0.25)."""

    def __init__(self, members, *, beta=1.754e-11, nu4=1e9, F0=2e-12, dt=1200.0, n_pre=3,
                 n_post=3, n_coarse=4, omega=0.85, rtol=1e-8, inner_rtol=1e-4, chebyshev=0.25):
        dev = _lib.require_gpu()
        self.members = list(members)
        self.n = len(self.members)
        self.beta, self.nu4, self.F0, self.dt = beta, nu4, F0, dt
        self.rtol, self.inner_rtol = rtol, inner_rtol
        shapes = [(m.ny + 2, m.nx + 2) for m in self.members]
        sizes = [a * b for a, b in shapes]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        total = int(self.offsets[-1])
        z = lambda: torch.zeros(total, dtype=torch.float64, device=dev)  # noqa: E731
        self.q, self.q1, self.q2 = z(), z(), z()
        self.ps = [z() for _ in range(4)]
        self.backup = z()
        self.masks = []
        self.solvers = []
        for k, m in enumerate(self.members):
            wet = np.ones((m.ny, m.nx)) if m.mask is None else np.asarray(m.mask, np.float64)
            self.solvers.append(build_multigrid_solver(
                wet, m.dx, m.dx, 1.0 / m.Ld**2, n_levels=hierarchy_depth(m.ny, m.nx, m.dx, m.Ld),
                n_pre=n_pre, n_post=n_post, n_coarse=n_coarse, omega=omega, kind="dirichlet",
                precision="mixed", chebyshev=chebyshev))
            if m.mask is None:
                self.masks.append(None)
            else:
                full = np.zeros(shapes[k], dtype=np.uint8)
                full[1:-1, 1:-1] = np.asarray(m.mask) > 0
                self.masks.append(torch.from_numpy(full).to(dev))
            qi = torch.tensor(m.q0 * (1.0 if m.mask is None else (np.asarray(m.mask) > 0)))
            self.view(self.q, k)[1:-1, 1:-1] = qi.to(dev)
        self.batch = MultigridBatch(self.solvers)
        self._stage = {}
        self._init_budgets()

    # -- buffers ---------------------------------------------------------
    def view(self, buf, k):
        m = self.members[k]
        o = int(self.offsets[k])
        return buf[o:o + (m.ny + 2) * (m.nx + 2)].view(m.ny + 2, m.nx + 2)

    def state(self, k):
        """Member k's full PV array (a view of the ensemble buffer)."""
        return self.view(self.q, k)

    def psi(self, k):
        """Member k's streamfunction of the last stage."""
        return self.view(self.ps[2], k)

    # -- launches --------------------------------------------------------
    def _stage_list(self, k, e):
        key = (k, e)
        if key not in self._stage:
            qs = (self.q, self.q1, self.q2)[k]
            out = (self.q1, self.q2, self.q)[k]
            psi = self._psi_buffer(k, e)
            a, b, c = COEF[k]
            descs = (_StageDesc * self.n)()
            for i, m in enumerate(self.members):
                d = descs[i]
                d.qs, d.psi = self.view(qs, i).data_ptr(), self.view(psi, i).data_ptr()
                d.qn, d.out = self.view(self.q, i).data_ptr(), self.view(out, i).data_ptr()
                d.mask = None if self.masks[i] is None else self.masks[i].data_ptr()
                d.Ny, d.Nx = m.ny + 2, m.nx + 2
                d.dx = d.dy = m.dx
                d.beta, d.Ly, d.F0, d.nu4 = self.beta, m.ny * m.dx, self.F0, self.nu4
                d.a, d.b, d.cdt = a, b, c * self.dt
            h = _vp()
            _lib.call("fvx_qg_stage_batch_create", self.n, descs, 1, C.byref(h))
            self._stage[key] = h
        return self._stage[key]

    def _psi_buffer(self, k, e):
        return stage_psi_buffer(self.ps, k, e)

    def _solve(self, k, e):
        qs = (self.q, self.q1, self.q2)[k]
        x = self._psi_buffer(k, e)
        rhs = [self.view(qs, i)[1:-1, 1:-1] for i in range(self.n)]
        xs = [self.view(x, i)[1:-1, 1:-1] for i in range(self.n)]
        no, ni = self.budgets[k]
        self.batch.enqueue(rhs, xs, rtol=self.rtol, n_outer=no, n_inner=ni,
                           inner_rtol=self.inner_rtol, slot=k)

    def _enqueue_step(self, e, fresh=False):
        self.backup.copy_(self.q)
        for k in range(3):
            extrapolated_guess(self.ps, k, e, fresh)
            self._solve(k, e)
            _lib.call("fvx_qg_stage_batch_launch", self._stage_list(k, e), _lib.stream())

    def _restore(self):
        self.q.copy_(self.backup)

    def _state(self):
        return [self.q] + list(self.ps)

    def _build_lists(self):
        for e in (0, 1):
            for k in range(3):
                self._stage_list(k, e)


def config4_members(n_members: int = 96, seed: int = 0, dx: float = 5000.0,
                    shapes=((128, 128), (256, 256), (512, 512), (768, 1024), (1024, 1024))):
    """The batched-ensemble configuration: member shapes drawn by seed from
    {128^2, 256^2, 512^2, 768 x 1024, 1024^2}, Ld log-uniform in [2e4, 1e5] m,
    every other member masked with the basin-mask generator at its own size,
    its own seeded k^-3 PV (rms 1e-5)."""
    from finitevolx_amd import inputs

    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_members):
        ny, nx = shapes[int(rng.integers(len(shapes)))]
        Ld = float(np.exp(rng.uniform(np.log(2e4), np.log(1e5))))
        mask = inputs.basin_mask(ny, nx, 1000 + k) if k % 2 == 1 else None
        q0 = inputs.qg_initial_pv(ny, nx, rms=1e-5, seed=2000 + k)
        out.append(Member(ny, nx, dx, Ld, q0, mask))
    return out


def autotune(members, candidates=None, warmup: int = 8, steps: int = 10, **kw):
    """Pick the smoother schedule of a batched ensemble by measurement: for
    each candidate (n_pre, n_post, omega) build the ensemble, run ``warmup``
    steps (budget calibration and tightening) and time ``steps`` graph steps.
    Returns ``(best candidate, {candidate: steps/s})``."""
    candidates = candidates or [(2, 2, 0.85), (3, 3, 0.85), (3, 3, 0.8), (2, 2, 0.9)]
    rates = {}
    for c in candidates:
        ens = QGEnsemble(members, n_pre=c[0], n_post=c[1], omega=c[2], **kw)
        ens.step(warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = ens.step(steps)
        torch.cuda.synchronize()
        rates[c] = steps / (time.perf_counter() - t0) if info["converged"] else 0.0
        del ens
    best = max(rates, key=rates.get)
    return best, rates
