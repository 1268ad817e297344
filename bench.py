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


def cpu_baseline():
    """Reference step loop on the host cores (MPI ranks) over a bounded
    sample: 1024^2 vortex, 40 steps after one warm-up step."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    n, steps = 1024, 40
    ncores = min(16, len(os.sched_getaffinity(0)))
    mpiexec = shutil.which("mpiexec") or ("/opt/conda/bin/mpiexec" if os.path.exists("/opt/conda/bin/mpiexec")
                                          else None)
    sample = f"2D vortex {n}^2, WENO5+HLLC+RK3 FP64, {steps} steps (reference step loop, 1 warm-up step)"
    if os.path.exists(exe):
        tries = []
        if mpiexec and ncores > 1:
            tries.append(([mpiexec, "-n", str(ncores), exe, "vortex2d", str(n), str(steps), "hllc"], ncores))
        tries.append(([exe, "vortex2d", str(n), str(max(4, steps // 8)), "hllc"], 1))
        for cmd, cores in tries:
            try:
                r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=300)
                if r.returncode == 0:
                    line = [l for l in r.stdout.decode().splitlines() if l.startswith("{")][-1]
                    d = json.loads(line)
                    smp = sample if cores > 1 else sample.replace(f"{steps} steps", f"{d['steps']} steps")
                    return {"value": d["cell_updates_per_s"], "unit": "cell-updates/s", "cores": d["ranks"],
                            "kind": "reference", "sample": smp + f", {d['ranks']} MPI rank(s), {d['seconds']:.2f} s"}
            except Exception:
                continue
    # fallback: the oracle restatement, one core
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
            "sample": f"2D vortex 256^2, {k} steps, C restatement, 1 core, {el:.2f} s"}


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
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
                     "mean_launch_ms": mean_launch_s * 1e3, "launches": res["launches"]},
        "finite": res["finite"],
    }
    if not a.no_variants and a.precision == "fp64":
        import numpy as np

        m = run_gpu(a.n, a.steps, a.warmup, "mixed", a.rows)
        mv = cells * a.steps / m["seconds"]
        ref = res["rho"]
        l1 = float(np.abs(m["rho"] - ref).sum() / np.abs(ref).sum())
        linf = float(np.abs(m["rho"] - ref).max() / np.abs(ref).max())
        mb = 3 * nv * 4 * cells
        mm = m["kernel_ms"] / 1e3 / max(m["launches"], 1)
        out["variants"] = {"mixed": {"value": mv, "ratio_vs_fp64": mv / value, "rho_rel_L1_vs_fp64": l1,
                                     "rho_rel_Linf_vs_fp64": linf, "roofline_achieved_GBs": mb / mm / 1e9,
                                     "mean_launch_ms": mm * 1e3}}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
