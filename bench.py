"""Benchmark: QG 1.5-layer steps/s at 2048^2 on one MI355X.

Workload (one "step" = one rk3_ssp_step of the QG model = 3 Helmholtz solves
to 1e-8 relative FP64 residual + 3 fused tendency/stage kernels): closed
rectangular basin, Lx = Ly = 5.12e6 m, f0 = 9.375e-5, beta = 1.754e-11,
Ld = 4e4 m, nu4 = 1e9 m^4/s, F0 = 2e-12, dt = 1200 s, seeded k^-3 initial PV
(rms 1e-5).  Paths timed back to back on the same inputs:

  * ``fp64``  : FP64 state, stencils and multigrid-preconditioned CG
  * ``mixed`` : FP32 state and stencils, FP32 multigrid-PCG inner solves with
                FP64 iterative refinement of psi (the headline value)

Prints ONE JSON line (plus progress lines on stderr).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def time_steps(model, q0, steps, warmup, torch):
    q = model.new_state(q0)
    psi = torch.zeros(model.shape, dtype=torch.float64, device=q.device)
    infos = []
    for _ in range(warmup):
        q, psi, info = model.step(q, psi, 1)
        infos.append(info)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        q, psi, info = model.step(q, psi, 1)
        infos.append(info)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return q, psi, dt, infos[warmup:]


def stage_kernel_roofline(model, torch, reps=50):
    """Average duration of the fused QG stage kernel, timed with HIP events
    around a captured graph of ``reps`` launches (no host gaps)."""
    Ny, Nx = model.shape
    dev = torch.device("cuda")
    g = torch.Generator(device="cpu").manual_seed(0)
    qs = (torch.randn(Ny, Nx, generator=g, dtype=torch.float64) * 1e-5).to(dev, model.dtype)
    qn = qs.clone()
    psi = (torch.randn(Ny, Nx, generator=g, dtype=torch.float64) * 1e4).to(dev)
    out = torch.empty_like(qs)
    from finitevolx_amd import _lib

    def launch():
        _lib.call(f"fvx_qg_stage_{_lib.suffix(qs)}", qs.data_ptr(), psi.data_ptr(), qn.data_ptr(),
                  out.data_ptr(), None, Ny, Nx, model.dx, model.dy, model.beta, model.Ly,
                  model.F0, model.nu4, 0.75, 0.25, 0.25 * model.dt, _lib.stream())

    launch()
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            for _ in range(reps):
                launch()
    graph.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    n_replay = 5
    e0.record()
    for _ in range(n_replay):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    avg_s = e0.elapsed_time(e1) * 1e-3 / (n_replay * reps)
    es = 4 if model.dtype == torch.float32 else 8
    psi_b = 8
    # algorithmic bytes per cell: read q_s, psi (FP64), q_n, write q_out
    bytes_per_cell = 3 * es + psi_b
    algo = bytes_per_cell * Ny * Nx
    return avg_s, algo, bytes_per_cell


def cpu_baseline(n, q0, budget_s=30.0):
    """The oracle's FP64 QG step (NumPy; V(2,2)-preconditioned flexible CG to
    1e-8) timed on this host for a bounded number of steps."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import fvx

    p = fvx.QGParams(n, n, 5.12e6, 5.12e6)
    F = p.forcing()
    wet = np.ones((n, n), bool)
    levels = fvx.build_multigrid(wet.astype(np.float64), p.dx, p.dy, p.lambda_, kind="dirichlet")
    psi_prev = [None]

    def solver(rhs):
        x, it, res, ok = fvx.pcg(lambda u: fvx.apply_operator(u, levels[0]), rhs,
                                 x0=psi_prev[0],
                                 precond=lambda r: fvx.v_cycle(levels, np.zeros_like(r), r, 0,
                                                               2, 2, 20, 0.8),
                                 rtol=1e-8, atol=0.0, max_steps=100)
        psi_prev[0] = x
        return x

    q = np.zeros((n + 2, n + 2))
    q[1:-1, 1:-1] = q0
    t0 = time.perf_counter()
    steps = 0
    while True:
        q = fvx.qg_step(q, p, F, wet, solver)
        steps += 1
        if time.perf_counter() - t0 > budget_s / 2 or steps >= 3:
            break
    dt = time.perf_counter() - t0
    return steps / dt, steps, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--paths", default="fp64,mixed")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the QG path "
              "does not shard across GPUs in this build.", file=sys.stderr)
        sys.exit(2)

    import torch

    from finitevolx_amd import inputs
    from finitevolx_amd.qg import QGModel

    n = args.n
    L = 5.12e6
    q0 = inputs.qg_initial_pv(n, n, rms=1e-5, seed=2)
    results = {}
    finals = {}
    for path in args.paths.split(","):
        prec = {"fp64": "float64", "mixed": "mixed"}[path]
        model = QGModel(n, n, L, L, precision=prec)
        log(f"path {path}: warmup {args.warmup}, timing {args.steps} steps at {n}^2")
        q, psi, dt, infos = time_steps(model, q0, args.steps, args.warmup, torch)
        sps = args.steps / dt
        it = [i["iterations"] for i in infos]
        results[path] = dict(
            steps_per_s=sps, ms_per_step=1e3 * dt / args.steps,
            pcg_iters_per_solve=float(np.sum(it)) / max(1, sum(i["solves"] for i in infos)),
            outer_per_solve=float(sum(i["outer"] for i in infos)) / max(1, sum(i["solves"] for i in infos)),
            max_rel_residual=max(i["max_rel_residual"] for i in infos),
            converged=all(i["converged"] for i in infos))
        finals[path] = q.double().cpu().numpy()
        log(f"path {path}: {sps:.2f} steps/s, {results[path]}")
        if path == args.paths.split(",")[-1]:
            avg_s, algo, bpc = stage_kernel_roofline(model, torch)
            achieved = algo / avg_s / 1e9
            roof = {"kernel": f"qg_stage_kernel<{'float' if model.dtype == torch.float32 else 'double'},double>",
                    "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": None, "avg_us": round(avg_s * 1e6, 2),
                    "algorithmic_bytes_per_cell": bpc}
            log(f"roofline: {roof}")
        del model
    head = "mixed" if "mixed" in results else list(results)[0]
    pv_err = None
    if "fp64" in finals and "mixed" in finals:
        ref = finals["fp64"]
        pv_err = float(np.max(np.abs(finals["mixed"] - ref)) / np.max(np.abs(ref)))
    cpu = None
    if not args.no_cpu:
        log("cpu baseline: oracle FP64 QG step (NumPy) ...")
        sps_cpu, nst, tcpu = cpu_baseline(n, q0)
        cpu = {"value": round(sps_cpu, 5), "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"{nst} oracle QG step(s) at {n}^2 (NumPy FP64, V(2,2) flexible PCG "
                         f"to 1e-8), {tcpu:.1f} s; {len(os.sched_getaffinity(0))} cores available"}
        log(f"cpu baseline: {cpu}")
    r = results[head]
    line = {
        "metric": "QG steps/s at 2048^2 (learned-PCG vs MG-PCG, rel resid 1e-8) + HBM GB/s",
        "value": round(r["steps_per_s"], 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 stencils/inner solves + f64 refinement" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"QG 1.5-layer closed basin {n}x{n}, rk3_ssp_step dt=1200 s, "
                               "Helmholtz to 1e-8 FP64 relative residual per stage",
                   "grid": [n, n], "path": head},
        "roofline": roof,
        "cpu_baseline": cpu,
        "paths": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                  for k, v in results.items()},
        "speedup_mixed_vs_fp64": (round(results["mixed"]["steps_per_s"] / results["fp64"]["steps_per_s"], 3)
                                  if "mixed" in results and "fp64" in results else None),
        "pv_rel_err_mixed_vs_fp64": pv_err,
        "cell_updates_per_s": round(r["steps_per_s"] * n * n, 1),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
