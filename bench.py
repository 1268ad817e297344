"""Benchmark: WENO5 + HLLC + SSP-RK3 on the 2D isentropic vortex.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 10] [--n 4096]
                    [--precision fp64|mixed] [--no-cpu] [--no-variants]

Workload: periodic [0,10]^2 vortex (rho=p=u=v=1 freestream, beta=5 at
(5,5), gamma=1.4), n x n cells, WENO5 + HLLC + SSP-RK3, CFL 0.6.  One step
is one full RK3 time step (3 fused stage kernels + the dt finalize) on the
device-resident state; the timed region runs K steps through the
device-side time loop (sifv_explicit_advance) between two synchronisations.

value = interior cell-updates per second (cells x steps / seconds).
roofline: algorithmic HBM bytes of one fused stage launch = 3 * nv * w per
cell (read U, write U, read or write the RK backup; nv = 4, w = 8 for FP64)
over the launch's mean duration from HIP events recorded around every
stage launch on the library's stream.
cpu_baseline: the reference itself (oracle/_ref/ref_bench, built from the
reference sources) timed on this host's cores through MPI ranks on a
bounded sample of the same workload; falls back to the oracle restatement.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--variants", default="mixed,vortex1024,ensemble,adjoint,bubble,graphs,igr",
                    help="comma list of the side workloads to run with an FP64 main run")
    ap.add_argument("--bubble-n", type=int, default=512, help="cells per side of the 3D bubble workload")
    ap.add_argument("--bubble-only", action="store_true",
                    help="only run the 3D bubble (FP64) for --steps steps and print its line (profiling the 3D "
                         "two-phase stage kernels and the pressure solvers)")
    ap.add_argument("--bubble-solver", default="jacobi", choices=["jacobi", "pcg", "mgpcg"])
    ap.add_argument("--bubble-precision", default="fp64", choices=["fp64", "mixed"],
                    help="mixed: FP32 Jacobi sweeps + FP64 refinement, or the FP32 V-cycle of the multigrid PCG")
    ap.add_argument("--bubble-stirred", action="store_true", help="the stirred bubble (stiff pressure solves)")
    ap.add_argument("--stage-slabs", type=int, default=0,
                    help="3D bubble: slabs of z segments per stage (0: the library's choice, 1: one launch per pass)")
    ap.add_argument("--bubble-graphs", action="store_true",
                    help="replay the Jacobi sweep chunks from captured HIP graphs")
    ap.add_argument("--igr-only", action="store_true",
                    help="only run the IGR side line (2D vortex, periodic and outflow faces, 1024^2 and 2048^2) and "
                         "print it (profiling the IGR kernels)")
    ap.add_argument("--rows", type=int, default=0, help="rows marched per wave (0 = built-in choice)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (HBM counter calibration) and exit")
    return ap.parse_args()


def copy_probe(mib: int, reps: int = 5):
    """Plain device-to-device copies of a known size through the HIP runtime's
    copy kernel: the reference point for scaling FETCH_SIZE / WRITE_SIZE."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": n, "reps": reps}))


def accuracy_1024(steps: int = 1000):
    """FP32-bulk against FP64 on the 1024^2 vortex after `steps` steps: rel L1
    and Linf of rho, rhoU, rhoV, rhoE (the FP64 run is within 1e-12 of the
    FP64 oracle by the parity tests)."""
    import numpy as np

    from semiimplicitfv_amd import ExplicitSolver, cases

    fields = {}
    for prec in ("fp64", "mixed"):
        cfg, mesh, st, eos, rs = cases.isentropic_vortex_2d(1024)
        g = ExplicitSolver(mesh, rs, eos, None, cfg, precision=prec)
        g.upload(st)
        g.advance(cfg, steps, 1e30, 0.0)
        fields[prec] = {f: mesh.interior(g.download_field(f)).copy() for f in (0, 1, 2, 4)}
        g.close()
    out = {"steps": steps}
    for f, name in ((0, "rho"), (1, "rhoU"), (2, "rhoV"), (4, "rhoE")):
        a, b = fields["mixed"][f], fields["fp64"][f]
        out[name] = {"rel_L1": float(np.abs(a - b).sum() / np.abs(b).sum()),
                     "rel_Linf": float(np.abs(a - b).max() / np.abs(b).max())}
    return out


def run_gpu(n, steps, warmup, precision, rows=0):
    import numpy as np

    from semiimplicitfv_amd import ExplicitSolver, cases

    cfg, mesh, st, eos, rs = cases.isentropic_vortex_2d(n)
    g = ExplicitSolver(mesh, rs, eos, None, cfg, precision=precision)
    if rows:
        g.set_tile(rows)
    g.upload(st)
    if warmup:
        g.advance(cfg, warmup, 1e30, 0.0)
    g.synchronize()
    g.timing(True)
    g.read_timing(reset=True)
    t0 = time.perf_counter()
    t_sim, done = g.advance(cfg, steps, 1e30, 0.0)
    g.synchronize()
    el = time.perf_counter() - t0
    kms, launches = g.read_timing(reset=True)
    g.timing(False)
    assert done == steps, (done, steps)
    rho = mesh.interior(g.download_field(0)).copy()
    finite = bool(np.isfinite(rho).all())
    g.close()
    return {"seconds": el, "kernel_ms": kms, "launches": launches, "rho": rho, "finite": finite,
            "cells": n * n, "dt_sum": t_sim}


def bench_graphs(sizes=(128, 256, 512), steps: int = 2000):
    """Step replay from a captured HIP graph vs one enqueue per launch
    (the library default), on small 2D vortex grids where per-step
    submission cost would matter most."""
    from semiimplicitfv_amd import ExplicitSolver, cases

    out = {}
    for n in sizes:
        cfg, mesh, st, eos, rs = cases.isentropic_vortex_2d(n)
        row = {}
        for graphs in (True, False):
            g = ExplicitSolver(mesh, rs, eos, None, cfg)
            g.set_graphs(graphs)
            g.upload(st)
            g.advance(cfg, 20, 1e30, 0.0)
            g.synchronize()
            t0 = time.perf_counter()
            g.advance(cfg, steps, 1e30, 0.0)
            g.synchronize()
            el = time.perf_counter() - t0
            row["graph_us_per_step" if graphs else "launch_us_per_step"] = el / steps * 1e6
            g.close()
        row["speedup"] = row["launch_us_per_step"] / row["graph_us_per_step"]
        out[f"vortex{n}"] = row
    return out


def bench_igr(sizes=(1024, 2048), steps: int = 20):
    """2D vortex with IGR (LF + WENO5 + SSP-RK3, IGRIters 5 exact
    Gauss-Seidel sweeps per stage, alphaCoeff 1) against the same run
    without IGR: wall time per step after a warm-up step (whose 50 warm-start
    sweeps are excluded).  Rows: the periodic vortex ("vortex{n}") and the same
    state between outflow faces ("outflow{n}"); a periodic x axis makes the
    strip-streaming sweep wait once per sweep for the strips' skew around the
    axis."""
    from semiimplicitfv_amd import ExplicitSolver, cases

    out = {}
    for n, periodic in [(n, p) for n in sizes for p in (True, False)]:
        row = {}
        for igr in (False, True):
            cfg, mesh, st, eos, rs = cases.isentropic_vortex_2d(n, riemann="lf", periodic=periodic)
            if igr:
                cfg.useIGR = True
                cfg.igrParams.alphaCoeff = 1.0
                cfg.igrParams.IGRIters = 5
                cfg.igrParams.IGRWarmStartIters = 50
                cfg.validate()
            g = ExplicitSolver(mesh, rs, eos, None, cfg)
            g.upload(st)
            t, _ = g.advance(cfg, 1, 1e30, 0.0)
            g.synchronize()
            cfg.step = 1
            t0 = time.perf_counter()
            g.advance(cfg, steps, 1e30, t)
            g.synchronize()
            el = time.perf_counter() - t0
            g.close()
            row["igr_ms_per_step" if igr else "plain_ms_per_step"] = el / steps * 1e3
        row["igr_sweeps_per_step"] = 15
        row["ms_per_igr_sweep"] = (row["igr_ms_per_step"] - row["plain_ms_per_step"]) / 15
        row["cell_sweeps_per_s"] = n * n / (row["ms_per_igr_sweep"] * 1e-3)
        out[f"{'vortex' if periodic else 'outflow'}{n}"] = row
    return out


def bench_ensemble(ntubes: int = 4096):
    """Batched Riemann ensemble: 4096 shock tubes (cells uniform over
    {256, ..., 4096}, states from cases.riemann_ensemble_params), WENO5 +
    HLLC + SSP-RK3, CFL 0.6, each to t = 0.1, in one persistent launch per
    run; tile shape from the autotuner (trials capped at 30 steps per tube).
    FP64, then FP32 bulk / FP64 EOS with its density error against FP64."""
    import numpy as np

    from semiimplicitfv_amd import RiemannEnsemble, _abi, cases, ensemble

    t0 = time.perf_counter()
    p = cases.riemann_ensemble_params(ntubes)
    cfg, eos, rs, tubes = ensemble.build_tubes(p)
    build_s = time.perf_counter() - t0
    out = {"workload": f"{ntubes} 1D shock tubes, {int(p['ncells'].sum())} cells, seed {cases.ENSEMBLE_SEED}, "
                       "WENO5+HLLC+SSP-RK3, CFL 0.6, t_end 0.1", "host_setup_s": build_s,
           "note": "tubes_non_finite: near-vacuum Riemann problems of the table on which the scheme itself fails "
                   "(the compiled reference reaches negative density on the same tubes); FP32 errors are over "
                   "tubes finite in both runs"}
    rho64 = None
    for prec in ("fp64", "mixed"):
        ens = RiemannEnsemble(tubes, cfg, rs, eos, precision=prec)
        block, wg, tune_ms = ens.autotune(0.1, 30)
        steps, times, ms = ens.run(0.1)
        cu = float((steps.astype(np.float64) * p["ncells"]).sum())
        rho = ens.download(_abi.F_RHO)
        ens.close()
        r = {"value": cu / (ms / 1e3), "unit": "cell-updates/s", "kernel_ms": ms, "cell_updates": cu,
             "steps_min": int(steps.min()), "steps_max": int(steps.max()), "all_reached_t_end": bool(
                 np.all(np.abs(times - 0.1) <= 1e-12)), "tile": {"block": block, "wg_per_cu": wg}}
        finite = np.array([bool(np.isfinite(a).all()) for a in rho])
        r["tubes_non_finite"] = int((~finite).sum())
        if prec == "fp64":
            rho64, fin64 = rho, finite
        else:
            both = finite & fin64
            errs = np.array([float(np.abs(a - b).sum() / np.abs(b).sum()) for a, b, ok in zip(rho, rho64, both) if ok])
            r["rho_rel_L1_vs_fp64_max_tube"] = float(errs.max()) if errs.size else None
            r["rho_rel_L1_vs_fp64_median"] = float(np.median(errs)) if errs.size else None
        out[prec] = r
    out["ratio_fp32_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    return out


def bench_adjoint(n: int = 512, nsteps: int = 400, every: int = 20, ndir: int = 64, h: float = 1e-6):
    """Differentiable run: 512^2 vortex with 8 seeded Gaussian density bumps
    (amplitude 1e-3, width 0.5), WENO5 + Rusanov + SSP-RK3 FP64, fixed dt
    (CFL 0.6 dt of the initial state), 400 steps, checkpoints every 20.
    J = sum 0.5 |rho u|^2 / rho dV of the final state.  Discrete adjoint
    gradient w.r.t. the initial conservative fields and gamma, checked
    against central differences (h = 1e-6) of GPU forward runs on 64 seeded
    random directions."""
    import numpy as np

    from semiimplicitfv_amd import ExplicitSolver, IdealGasEOS, RusanovSolver, _abi, cases

    cfg, mesh, st, eos, rs, _ = cases.perturbed_vortex_2d(n)
    F = (_abi.F_RHO, _abi.F_RHOU, _abi.F_RHOV, _abi.F_RHOE)
    I = mesh.interior
    dV = (10.0 / n) ** 2

    def objective(fields):
        rho, mu, mv = I(fields[0]), I(fields[1]), I(fields[2])
        return float((0.5 * (mu * mu + mv * mv) / rho * dV).sum())

    def solver(gamma):
        e = IdealGasEOS(gamma, 1.0, cfg)
        return ExplicitSolver(mesh, RusanovSolver(e), e, None, cfg)

    g = solver(1.4)
    g.upload(st)
    g.synchronize()
    t0 = time.perf_counter()
    g.adjoint_forward(cfg, nsteps, every)
    t_fwd = time.perf_counter() - t0
    final = [g.download_field(f) for f in F]
    J = objective(final)
    seed = [np.zeros(mesh.totalCells()) for _ in range(4)]
    rho, mu, mv = I(final[0]), I(final[1]), I(final[2])
    I(seed[0])[...] = -0.5 * (mu * mu + mv * mv) / (rho * rho) * dV
    I(seed[1])[...] = mu / rho * dV
    I(seed[2])[...] = mv / rho * dV
    t0 = time.perf_counter()
    grad, dgam = g.adjoint_backward(seed)
    t_bwd = time.perf_counter() - t0
    g.close()

    def J_of(fields0, gamma):
        s = solver(gamma)
        for f, a in zip(F, fields0):
            s.upload_field(f, a)
        s.adjoint_forward(cfg, nsteps, nsteps)
        v = objective([s.download_field(f) for f in F])
        s.close()
        return v

    rng = np.random.default_rng(cases.ADJOINT_SEED + 1)
    base = [st.field(f).copy() for f in F]
    errs, abs_errs, dots, errs4 = [], [], [], []
    h4 = 1e-4

    def J_at(d, dg, t):
        return J_of([a + t * b for a, b in zip(base, d)], 1.4 + t * dg)

    for _ in range(ndir):
        d = []
        for a in base:
            x = np.zeros_like(a)
            I(x)[...] = rng.standard_normal(I(x).shape) * 1e-2 * np.abs(I(a)).max()
            d.append(x)
        dg = 0.05 * rng.standard_normal()
        fd = (J_at(d, dg, h) - J_at(d, dg, -h)) / (2 * h)
        # fourth-order central difference at a larger step: truncation O(h^4),
        # rounding noise of J over 2h 100x smaller than at h = 1e-6
        fd4 = (8.0 * (J_at(d, dg, h4) - J_at(d, dg, -h4)) - (J_at(d, dg, 2 * h4) - J_at(d, dg, -2 * h4))) / (12 * h4)
        dot = sum(float((I(gq) * I(dq)).sum()) for gq, dq in zip(grad, d)) + dgam * dg
        errs.append(abs(dot - fd) / abs(fd))
        abs_errs.append(abs(dot - fd))
        errs4.append(abs(dot - fd4) / abs(fd4))
        dots.append(abs(dot))
    e, ea, e4 = np.array(errs), np.array(abs_errs), np.array(errs4)
    return {"workload": f"{n}^2 perturbed vortex, WENO5+Rusanov+SSP-RK3 FP64, {nsteps} steps, fixed dt "
                        f"{cfg.explicitParams.constDt:.4e}, checkpoint every {every}",
            "J": J, "dJ_dgamma": dgam, "forward_s": t_fwd, "backward_s": t_bwd,
            "fd_directions": ndir, "fd_h": h, "fd_rel_err_median": float(np.median(e)),
            "fd_rel_err_p90": float(np.percentile(e, 90)), "fd_rel_err_max": float(e.max()),
            "fd_within_1e-6": int((e <= 1e-6).sum()),
            "fd_abs_err_max": float(ea.max()), "median_abs_directional_derivative": float(np.median(dots)),
            "min_abs_directional_derivative": float(np.min(dots)),
            "fd4_h": h4, "fd4_rel_err_median": float(np.median(e4)), "fd4_rel_err_max": float(e4.max()),
            "fd4_within_1e-6": int((e4 <= 1e-6).sum()),
            "note": ("central differences at h = 1e-6 carry the rounding noise of J (~1e-14 on J ~ 100) over "
                     "2h, ~3e-8 absolute: every direction's |adjoint - FD| sits at that floor, and the relative "
                     "misses are the directions whose derivative nearly cancels (min |dJ.d| / median). The "
                     "fourth-order difference at h = 1e-4 removes that floor (fd4_*).")}


def bench_bubble(n: int = 512, steps: int = 3, precision: str = "fp64", stirred: bool = False,
                 solver: str = "jacobi", graphs: bool = False, slabs: int = 0):
    """3D water/air bubble (unit cube, n^3 cells, air sphere r = 0.2, alpha
    tanh over 2 cells, alphaMin 1e-8, slip walls, p = 1e5 at rest), semi-
    implicit SSP-RK3 with the Jacobi pressure solve (tol 1e-6, <= 1000
    sweeps), WENO5 + HLLC, FP64.  Fields are built one at a time to bound host
    memory.  One step = 3 advective stages + 3 pressure solves + corrections.
    precision "mixed": the same FP64 state with the pressure solve run as FP32
    Jacobi sweeps refined by FP64 residual passes (solver "jacobi") or with
    the multigrid V-cycle in FP32 inside FP64 conjugate gradients (solver
    "mgpcg").  stirred: the velocity
    field of the 12^3 bubble fixture (u = 2 sin(pi x) cos(pi y), v = -2 cos(pi x)
    sin(pi y), w = 0.5 sin(pi z)), which makes every pressure solve run to its
    1000-sweep cap.  solver "pcg": the conjugate-gradient pressure solve
    (stop on ||rhs - A p||_2 <= 1e-6 ||rhs||_2, <= 1000 iterations) instead of
    Jacobi (stop on max |p_new - p| < 1e-6); solver "mgpcg": the same with the
    multigrid V-cycle preconditioner.  graphs: Jacobi chunks replayed from
    captured HIP graphs (captured in the warm-up step) instead of enqueued
    sweep by sweep.  slabs: slabs of z segments per 3D stage (0: the library's
    choice)."""
    import numpy as np

    from semiimplicitfv_amd import (BoundaryCondition, HLLCSolver, JacobiPressureSolver, MGPCGPressureSolver,
                                    PCGPressureSolver, PhaseEOS, ReconstructionOrder, SemiImplicitSolver, SimulationConfig,
                                    StiffenedGasEOS, _abi, uniform_mesh)

    WATER, AIR = PhaseEOS(4.4, 6.0e8), PhaseEOS(1.4, 0.0)

    cfg = SimulationConfig(dim=3, nGhost=3, RKOrder=3, reconOrder=ReconstructionOrder.WENO5, semiImplicit=True)
    cfg.semiImplicitParams.maxDt = 1e-2
    cfg.semiImplicitParams.pressureTol = 1e-6
    cfg.semiImplicitParams.maxPressureIters = 1000
    cfg.multiPhaseParams.nPhases = 2
    cfg.multiPhaseParams.phases = [WATER, AIR]
    cfg.multiPhaseParams.alphaMin = 1e-8
    cfg.validate()
    mesh = uniform_mesh(cfg, n, 0.0, 1.0, n, 0.0, 1.0, n, 0.0, 1.0)
    for f in range(6):
        mesh.setBoundaryCondition(f, BoundaryCondition.SlipWall)
    eos = StiffenedGasEOS(4.4, 6.0e8, 1.0, cfg)
    t0 = time.perf_counter()
    ps = {"pcg": PCGPressureSolver, "mgpcg": MGPCGPressureSolver}.get(solver, JacobiPressureSolver)()
    g = SemiImplicitSolver(mesh, HLLCSolver(eos), ps, eos, None, cfg, precision=precision)
    g.set_graphs(graphs)
    if slabs:
        g.set_stage_slabs(slabs)
    xc, yc, zc = mesh.centroids()
    r = np.sqrt((xc[None, None, :] - 0.5) ** 2 + (yc[None, :, None] - 0.5) ** 2 + (zc[:, None, None] - 0.5) ** 2)
    amin = cfg.multiPhaseParams.alphaMin
    aw = np.clip(0.5 * (1.0 + np.tanh((r - 0.2) / (2.0 / n))), amin, 1.0 - amin)
    del r
    p0 = 1e5
    fields = {
        _abi.F_ALPHA: lambda: aw,
        _abi.F_ALPHA + 1: lambda: 1.0 - aw,
        _abi.F_ALPHARHO: lambda: aw * 1000.0,
        _abi.F_ALPHARHO + 1: lambda: (1.0 - aw) * 1.0,
        _abi.F_RHO: lambda: aw * 1000.0 + (1.0 - aw) * 1.0,
        _abi.F_PRES: lambda: np.full(aw.shape, p0),
        _abi.F_RHOE: lambda: aw * (p0 + 4.4 * 6.0e8) / 3.4 + (1.0 - aw) * p0 / 0.4,
    }
    if stirred:
        X, Y, Z = xc[None, None, :], yc[None, :, None], zc[:, None, None]
        u = (2.0 * np.sin(np.pi * X) * np.cos(np.pi * Y), -2.0 * np.cos(np.pi * X) * np.sin(np.pi * Y),
             0.5 * np.sin(np.pi * Z))
        rho_f = lambda: aw * 1000.0 + (1.0 - aw) * 1.0  # noqa: E731
        ke = lambda: 0.5 * rho_f() * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)  # noqa: E731
        for d, (mf, vf) in enumerate(((_abi.F_RHOU, _abi.F_VELU), (_abi.F_RHOV, _abi.F_VELV), (_abi.F_RHOW, _abi.F_VELW))):
            fields[mf] = (lambda d=d: rho_f() * u[d] * np.ones_like(aw))
            fields[vf] = (lambda d=d: u[d] * np.ones_like(aw))
        e0 = fields[_abi.F_RHOE]
        fields[_abi.F_RHOE] = lambda: e0() + ke()
    buf = np.zeros(mesh.totalCells())
    for fid, make in fields.items():
        buf[...] = 0.0
        mesh.interior(buf)[...] = make()
        g.upload_field(fid, buf)
    del buf, aw
    setup_s = time.perf_counter() - t0
    cfg.step = 0
    g.step_resident(cfg)  # warm-up step
    g.synchronize()
    t0 = time.perf_counter()
    iters = []
    for k in range(steps):
        cfg.step = k + 1
        g.step_resident(cfg)
        iters.append(g.lastPressureIterations())
    g.synchronize()
    el = time.perf_counter() - t0
    pres = mesh.interior(g.download_field(_abi.F_PRES)).copy()
    finite = bool(np.isfinite(pres).all())
    g.close()
    solve = {"pcg": "PCG (relative residual 1e-6)",
             "mgpcg": "multigrid-preconditioned PCG (relative residual 1e-6)"}.get(solver, "Jacobi (tol 1e-6)")
    if precision != "mixed":
        prec = "throughout"
    elif solver == "mgpcg":
        prec = "state and conjugate gradients, FP32 V-cycle"
    else:
        prec = "state, FP32 Jacobi sweeps + FP64 refinement"
    return {"workload": f"3D water/air bubble {n}^3, semi-implicit SSP-RK3 + {solve}, WENO5+HLLC, FP64 " + prec,
            "value": n ** 3 * steps / el, "unit": "cell-updates/s", "ms_per_step": el / steps * 1e3,
            "pressure_iterations_last_stage": iters, "host_setup_s": setup_s, "finite": finite, "_pres": pres}


def measured_traffic(n, precision):
    """HBM bytes per stage launch from the committed rocprofv3 PMC summary of
    this workload (FETCH_SIZE + WRITE_SIZE scaled by a copy-kernel
    calibration), or None when no summary exists for it."""
    name = f"stage_xy_weno5_hllc_{precision}_vortex{n}_pmc.md"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


# Issue model of one wave64 VALU instruction on MI355X: 4 cycles of a 16-lane
# SIMD (FP64 and unpacked FP32 alike), 256 CUs x 4 SIMDs, 2.4 GHz peak clock.
VALU_SIMDS = 1024
VALU_CYCLES_PER_INST = 4
CLOCK_HZ = 2.4e9


def measured_valu(n, precision, mean_launch_s):
    """VALU issue time per stage launch from the committed counter summary
    (SQ_INSTS_VALU wave-instructions per dispatch) against the measured launch
    time: the fraction of the launch the SIMDs spend issuing VALU work."""
    name = f"stage_xy_weno5_hllc_{precision}_vortex{n}_pmc.md"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("valu_wave_instructions_per_launch:"):
            insts = float(line.split(":", 1)[1])
            issue_s = insts * VALU_CYCLES_PER_INST / (VALU_SIMDS * CLOCK_HZ)
            return {"wave_instructions": insts, "issue_ms": issue_s * 1e3,
                    "frac": issue_s / mean_launch_s if mean_launch_s > 0 else None,
                    "note": "VALU issue time / measured launch time; the FP64 stage kernel is VALU-bound"}
    return None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(n: int = 4096, steps: int = 3):
    """Reference step loop on the host cores (MPI ranks) over a bounded
    sample of the GPU line's own workload: the n^2 vortex, `steps` steps
    after one warm-up step.  Ranks: every core this process may use, capped
    at 16, the host-core share that goes with one GPU on the MI355X boxes
    (their process guard and OMP_NUM_THREADS=16 set that share)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    avail = len(os.sched_getaffinity(0))
    ncores = min(16, avail)
    mpiexec = shutil.which("mpiexec") or ("/opt/conda/bin/mpiexec" if os.path.exists("/opt/conda/bin/mpiexec")
                                          else None)
    model = cpu_model()
    sample = (f"2D vortex {n}^2 (the GPU line's workload), WENO5+HLLC+RK3 FP64, {steps} steps "
              f"(reference step loop, 1 warm-up step), {model}, {ncores} of {avail} visible cores")
    if os.path.exists(exe):
        tries = []
        if mpiexec and ncores > 1:
            tries.append(([mpiexec, "-n", str(ncores), exe, "vortex2d", str(n), str(steps), "hllc"], ncores))
        tries.append(([exe, "vortex2d", str(min(n, 1024)), "4", "hllc"], 1))
        for cmd, cores in tries:
            try:
                r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=300)
                if r.returncode == 0:
                    line = [l for l in r.stdout.decode().splitlines() if l.startswith("{")][-1]
                    d = json.loads(line)
                    smp = sample if cores > 1 else (f"2D vortex {min(n, 1024)}^2, WENO5+HLLC+RK3 FP64, {d['steps']} "
                                                    f"steps (reference step loop), {model}, 1 core")
                    return {"value": d["cell_updates_per_s"], "unit": "cell-updates/s", "cores": d["ranks"],
                            "kind": "reference", "sample": smp + f", {d['ranks']} MPI rank(s), {d['seconds']:.2f} s"}
            except Exception:
                continue
    # fallback: the oracle restatement, one core (smaller sample: one core)
    from oracle import pyoracle
    from semiimplicitfv_amd import cases

    if not pyoracle.available("c"):
        return None
    cfg, mesh, st, eos, rs = cases.isentropic_vortex_2d(256)
    cpu = pyoracle.CpuSolver(mesh, cfg, rs.kind(), eos, "c")
    cpu.explicit_step(st, cfg)
    t0 = time.perf_counter()
    k = 10
    for _ in range(k):
        cpu.explicit_step(st, cfg)
    el = time.perf_counter() - t0
    return {"value": 256 * 256 * k / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"2D vortex 256^2, {k} steps, C restatement, 1 core of {model}, {el:.2f} s"}


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.bubble_only:
        r = bench_bubble(a.bubble_n, a.steps, precision=a.bubble_precision, stirred=a.bubble_stirred,
                         solver=a.bubble_solver, graphs=a.bubble_graphs, slabs=a.stage_slabs)
        r.pop("_pres", None)
        print(json.dumps(r))
        return
    if a.igr_only:
        print(json.dumps(bench_igr(steps=min(a.steps, 20))))
        return
    if a.gpus != 1:
        print(json.dumps({"error": "bench.py measures the path on one GPU; runs on several GPUs are not supported",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    res = run_gpu(a.n, a.steps, a.warmup, a.precision, a.rows)
    cells = res["cells"]
    value = cells * a.steps / res["seconds"]
    w = 8 if a.precision == "fp64" else 4
    nv = 4
    bytes_per_launch = 3 * nv * w * cells
    mean_launch_s = res["kernel_ms"] / 1e3 / max(res["launches"], 1)
    achieved = bytes_per_launch / mean_launch_s / 1e9
    out = {
        "metric": "WENO5+HLLC cell-updates/s & HBM % of peak, FP32/batched vs FP64 at stated L1",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": res["seconds"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32-bulk/f64-eos",
        "data": "synthetic (analytic isentropic vortex initial condition)",
        "config": {"workload": f"2D isentropic vortex {a.n}x{a.n}, periodic [0,10]^2, WENO5+HLLC+SSP-RK3, CFL 0.6",
                   "cells": cells, "precision": a.precision},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(a.n, a.precision),
                     "kernel": "stage_xy_kernel<2,3,HLLC>", "bytes_per_launch": bytes_per_launch,
                     "mean_launch_ms": mean_launch_s * 1e3, "launches": res["launches"],
                     "valu": measured_valu(a.n, a.precision, mean_launch_s)},
        "finite": res["finite"],
    }
    want = set() if a.no_variants or a.precision != "fp64" else set(a.variants.split(","))
    out["variants"] = {}
    if "mixed" in want:
        import numpy as np

        m = run_gpu(a.n, a.steps, a.warmup, "mixed", a.rows)
        mv = cells * a.steps / m["seconds"]
        ref = res["rho"]
        l1 = float(np.abs(m["rho"] - ref).sum() / np.abs(ref).sum())
        linf = float(np.abs(m["rho"] - ref).max() / np.abs(ref).max())
        mb = 3 * nv * 4 * cells
        mm = m["kernel_ms"] / 1e3 / max(m["launches"], 1)
        out["variants"]["mixed"] = {"value": mv, "ratio_vs_fp64": mv / value, "rho_rel_L1_vs_fp64": l1,
                                    "rho_rel_Linf_vs_fp64": linf, "roofline_achieved_GBs": mb / mm / 1e9,
                                    "mean_launch_ms": mm * 1e3,
                                    "accuracy_1024_1000_steps": accuracy_1024(),
                                    "note": "error against the FP64 GPU run, itself within 1e-12 of the FP64 oracle"}
    if "vortex1024" in want:
        t0 = time.perf_counter()
        r64 = run_gpu(1024, 400, 20, "fp64")
        r32 = run_gpu(1024, 400, 20, "mixed")
        v64 = r64["cells"] * 400 / r64["seconds"]
        v32 = r32["cells"] * 400 / r32["seconds"]
        out["variants"]["vortex1024"] = {
            "workload": "2D isentropic vortex 1024x1024, WENO5+HLLC+SSP-RK3, CFL 0.6, 400 steps",
            "fp64": {"value": v64, "mean_launch_ms": r64["kernel_ms"] / max(r64["launches"], 1),
                     "ratio_vs_4096": v64 / value},
            "mixed": {"value": v32, "mean_launch_ms": r32["kernel_ms"] / max(r32["launches"], 1),
                      "ratio_vs_fp64": v32 / v64},
            "note": ("1M cells give ~3k waves of 62 columns x 6 rows at one wave per SIMD slot: the 2-row warm-up "
                     "of each wave's y-window is 1/3 of its rows (1/8 at 4096^2, 16 rows)"),
            "wall_s": time.perf_counter() - t0}
    def bubble_pair():
        import numpy as np

        out = {}
        for tag, n, steps, stirred in (("at_rest", a.bubble_n, 3, False), ("stirred_256", 256, 1, True)):
            r64 = bench_bubble(n, steps, precision="fp64", stirred=stirred)
            r32 = bench_bubble(n, steps, precision="mixed", stirred=stirred)
            rcg = bench_bubble(n, steps, precision="fp64", stirred=stirred, solver="pcg")
            rmg = bench_bubble(n, steps, precision="fp64", stirred=stirred, solver="mgpcg")
            rmg32 = bench_bubble(n, steps, precision="mixed", stirred=stirred, solver="mgpcg")
            p64, p32, pcg, pmg = r64.pop("_pres"), r32.pop("_pres"), rcg.pop("_pres"), rmg.pop("_pres")
            pmg32 = rmg32.pop("_pres")
            rmg32["pres_rel_L1_vs_mgpcg"] = float(np.abs(pmg32 - pmg).sum() / np.abs(pmg).sum())
            rmg32["ratio_vs_mgpcg"] = rmg32["value"] / rmg["value"]
            rmg["pres_rel_L1_vs_jacobi"] = float(np.abs(pmg - p64).sum() / np.abs(p64).sum())
            rmg["pres_rel_L1_vs_pcg"] = float(np.abs(pmg - pcg).sum() / np.abs(pcg).sum())
            rmg["ratio_vs_jacobi"] = rmg["value"] / r64["value"]
            r32["pres_rel_L1_vs_fp64"] = float(np.abs(p32 - p64).sum() / np.abs(p64).sum())
            r32["ratio_vs_fp64"] = r32["value"] / r64["value"]
            rcg["pres_rel_L1_vs_jacobi"] = float(np.abs(pcg - p64).sum() / np.abs(p64).sum())
            rcg["ratio_vs_jacobi"] = rcg["value"] / r64["value"]
            if stirred:
                r32["note"] = ("both solves stop at the 1000-sweep cap unconverged, so pres_rel_L1_vs_fp64 "
                               "compares two different partial iterates, not FP32 rounding")
            elif np.array_equal(p32, p64):
                r32["note"] = ("every pressure solve ends within the FP64 opening chunk of 4 sweeps that the mixed "
                               "form shares with the FP64 solver, so both run the same sweeps (same bits); the FP32 "
                               "sweeps start on solves that outlast it")
            out[tag] = {"fp64": r64, "fp32_sweeps": r32, "pcg": rcg, "mgpcg": rmg, "mgpcg_fp32_vcycle": rmg32}
        return out

    for name, fn in (("ensemble", bench_ensemble), ("adjoint", bench_adjoint), ("bubble", bubble_pair),
                     ("graphs", bench_graphs), ("igr", bench_igr)):
        if name in want:
            t0 = time.perf_counter()
            try:
                out["variants"][name] = fn()
            except Exception as e:  # a side workload must not hide the main line
                out["variants"][name] = {"error": repr(e)}
            out["variants"][name]["wall_s"] = time.perf_counter() - t0
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.n)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
