"""Writes the golden fixtures of tests/golden/ from the compiled reference.

    make -C oracle && python tests/golden/make_golden.py [generator ...]

Every expected output here comes from the unmodified reference sources
(oracle/_ref/libsifv_ref.so, built from the reference checkout by
oracle/Makefile) on seeded inputs; the inputs are stored next to the
outputs.  The fixtures are data only (float64 arrays in .npz files).
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from semiimplicitfv_amd import (BoundaryCondition, IdealGasEOS, ReconstructionOrder, SimulationConfig,  # noqa: E402
                                SolutionState, StiffenedGasEOS, cases, uniform_mesh)
from semiimplicitfv_amd.cases import _fill  # noqa: E402
from semiimplicitfv_amd.timeloop import run_time_loop  # noqa: E402
from tests.golden.problems import (body_force_cases, bubble_3d, capillary_cases, jacobi_problem, liquid_gas_tube,  # noqa: E402
                                   random_ghost_case, ensemble_sample, igr_cases, recon_lines, riemann_pairs,
                                   two_phase_cases, viscous_cases)

OUT = os.path.dirname(os.path.abspath(__file__))


def save(name, **arrs):
    np.savez_compressed(os.path.join(OUT, name), **arrs)
    print("wrote", name, sum(a.nbytes for a in arrs.values()), "bytes raw")


def g_recon():
    out = {}
    for order in (1, 2, 4, 5):
        nx, ng, lines = recon_lines(order)
        L = np.zeros((len(lines), nx + 1))
        R = np.zeros((len(lines), nx + 1))
        for q, v in enumerate(lines):
            L[q], R[q] = po.recon_line_ref(order, 1e-6, nx, ng, v)
        out[f"v{order}"] = lines
        out[f"L{order}"] = L
        out[f"R{order}"] = R
    save("recon_kat.npz", **out)


def g_riemann():
    out = {}
    for key, (Ls, Rs, meta) in riemann_pairs().items():
        rtype, d, dim, incP, igr, nph = meta
        F = np.array([po.riemann_flux("ref", rtype, l, r, d, dim, incP, igr, nph) for l, r in zip(Ls, Rs)])
        out[key + "_L"] = Ls
        out[key + "_R"] = Rs
        out[key + "_F"] = F
    save("riemann_kat.npz", **out)


def g_sod():
    cfg, mesh, st, eos, rs = cases.sod_1d(1024)
    s = po.CpuSolver(mesh, cfg, rs.kind(), eos, "ref")
    dts = []
    t = 0.0
    while t < 0.2:
        dt = s.explicit_step(st, cfg, 0.2 - t)
        dts.append(dt)
        t += dt
        cfg.step += 1
    save("sod1024_final.npz", dts=np.array(dts), rho=st.rho, rhoU=st.rhoU, rhoE=st.rhoE, velU=st.velU,
         pres=st.pres)


def g_vortex():
    cfg, mesh, st, eos, rs = cases.isentropic_vortex_2d(64)
    s = po.CpuSolver(mesh, cfg, rs.kind(), eos, "ref")
    dts = []
    for k in range(50):
        cfg.step = k
        dts.append(s.explicit_step(st, cfg))
    save("vortex64_50steps.npz", dts=np.array(dts), rho=st.rho, rhoU=st.rhoU, rhoV=st.rhoV, rhoE=st.rhoE)


def g_ghost():
    out = {}
    for dim in (1, 2, 3):
        cfg, mesh, st, bcs, periods = random_ghost_case(dim)
        s = po.CpuSolver(mesh, cfg, 2, IdealGasEOS(1.4, 1.0, cfg), "ref")
        for vs in (0, 1, 2):
            t = st.copy()
            s.fill_ghosts(t, vs)
            for fid in t.field_ids():
                out[f"d{dim}_vs{vs}_f{fid}"] = t.field(fid)
        a = st.pres.copy()
        s.fill_scalar(a)
        out[f"d{dim}_scalar"] = a
        for fid in st.field_ids():
            out[f"d{dim}_in_f{fid}"] = st.field(fid)
    save("ghost_fill.npz", **out)


def g_jacobi():
    cfg, mesh, rho, rhoc2, rhs, p0, dt = jacobi_problem()
    s = po.CpuSolver(mesh, cfg, 2, IdealGasEOS(1.4, 1.0, cfg), "ref")
    out = {"rho": rho, "rhoc2": rhoc2, "rhs": rhs, "p0": p0, "dt": np.array([dt])}
    for tol, mi in ((1e-6, 1000), (1e-300, 7)):
        p = p0.copy()
        it = s.jacobi(rho, rhoc2, rhs, p, dt, tol, mi)
        out[f"p_tol{tol:g}"] = p
        out[f"it_tol{tol:g}"] = np.array([it])
    save("jacobi_3d.npz", **out)


def g_semi():
    cfg, mesh, st, eos, rs = bubble_3d(12, perturb=True)
    s = po.CpuSolver(mesh, cfg, rs.kind(), eos, "ref")
    its, dts = [], []
    for k in range(3):
        cfg.step = k
        dt, it = s.semi_step(st, cfg)
        its.append(it)
        dts.append(dt)
    out = {"dts": np.array(dts), "iters": np.array(its)}
    for fid in st.field_ids():
        out[f"f{fid}"] = st.field(fid)
    save("bubble3d_semi.npz", **out)


def g_liquid_gas():
    cfg, mesh, st, eos, rs = liquid_gas_tube(200)
    s = po.CpuSolver(mesh, cfg, rs.kind(), eos, "ref")
    dts = []
    for k in range(60):
        cfg.step = k
        dts.append(s.explicit_step(st, cfg))
    out = {"dts": np.array(dts)}
    for fid in st.field_ids():
        out[f"f{fid}"] = st.field(fid)
    save("liquid_gas_1d.npz", **out)


def _run_cases(table, name):
    out = {}
    for key, (build, rs, nsteps, semi) in table.items():
        cfg, mesh, st, eos = build()
        s = po.CpuSolver(mesh, cfg, rs, eos, "ref")
        dts, its = [], []
        for k in range(nsteps):
            cfg.step = k
            if semi:
                dt, it = s.semi_step(st, cfg)
                its.append(it)
            else:
                dt = s.explicit_step(st, cfg)
            dts.append(dt)
        out[f"{key}_dts"] = np.array(dts)
        out[f"{key}_iters"] = np.array(its, dtype=np.int64)
        for fid in st.field_ids():
            out[f"{key}_f{fid}"] = np.ascontiguousarray(mesh.interior(st.field(fid)))
    save(name, **out)


def g_two_phase():
    """Interior fields, time steps and Jacobi counts of the seeded water/air
    cases (explicit and semi-implicit, 1D-3D)."""
    _run_cases(two_phase_cases(), "two_phase_ref.npz")


def g_igr():
    """The same for the IGR cases (explicit 1D-3D, semi-implicit 2D)."""
    _run_cases(igr_cases(), "igr_ref.npz")


def g_body_force():
    """Body-force cases, config.time advanced per step (runTimeLoop,
    src/RKTimeStepping.cpp:258-266)."""
    out = {}
    for key, (build, rs, nsteps, semi) in body_force_cases().items():
        cfg, mesh, st, eos = build()
        s = po.CpuSolver(mesh, cfg, rs, eos, "ref")
        dts, its, t = [], [], 0.0
        for k in range(nsteps):
            cfg.step, cfg.time = k, t
            if semi:
                dt, it = s.semi_step(st, cfg)
                its.append(it)
            else:
                dt = s.explicit_step(st, cfg)
            dts.append(dt)
            t += dt
        out[f"{key}_dts"] = np.array(dts)
        out[f"{key}_iters"] = np.array(its, dtype=np.int64)
        for fid in st.field_ids():
            out[f"{key}_f{fid}"] = np.ascontiguousarray(mesh.interior(st.field(fid)))
    save("body_force_ref.npz", **out)


def g_viscous():
    """Viscous cases, config.time advanced per step."""
    out = {}
    for key, (build, rs, nsteps, semi) in viscous_cases().items():
        cfg, mesh, st, eos = build()
        s = po.CpuSolver(mesh, cfg, rs, eos, "ref")
        dts, its, t = [], [], 0.0
        for k in range(nsteps):
            cfg.step, cfg.time = k, t
            if semi:
                dt, it = s.semi_step(st, cfg)
                its.append(it)
            else:
                dt = s.explicit_step(st, cfg)
            dts.append(dt)
            t += dt
        out[f"{key}_dts"] = np.array(dts)
        out[f"{key}_iters"] = np.array(its, dtype=np.int64)
        for fid in st.field_ids():
            out[f"{key}_f{fid}"] = np.ascontiguousarray(mesh.interior(st.field(fid)))
    save("viscous_ref.npz", **out)


def g_capillary():
    """Surface-tension cases, config.time advanced per step."""
    out = {}
    for key, (build, rs, nsteps, semi) in capillary_cases().items():
        cfg, mesh, st, eos = build()
        s = po.CpuSolver(mesh, cfg, rs, eos, "ref")
        dts, its, t = [], [], 0.0
        for k in range(nsteps):
            cfg.step, cfg.time = k, t
            if semi:
                dt, it = s.semi_step(st, cfg)
                its.append(it)
            else:
                dt = s.explicit_step(st, cfg)
            dts.append(dt)
            t += dt
        out[f"{key}_dts"] = np.array(dts)
        out[f"{key}_iters"] = np.array(its, dtype=np.int64)
        for fid in st.field_ids():
            out[f"{key}_f{fid}"] = np.ascontiguousarray(mesh.interior(st.field(fid)))
    save("capillary_ref.npz", **out)


def g_ensemble():
    """Shock tubes of the batched ensemble sample run to t = 0.1 one by one
    (runTimeLoop dt rule): step counts, final times and interior fields."""
    from semiimplicitfv_amd import ensemble

    p, idx = ensemble_sample()
    cfg, eos, rs, tubes = ensemble.build_tubes(p, idx)
    out = {"rows": np.array(idx)}
    for k, (mesh, st) in enumerate(tubes):
        s = po.CpuSolver(mesh, cfg, rs.kind(), eos, "ref")
        t, n = 0.0, 0
        while t < 0.1:
            cfg.step = n
            t += s.explicit_step(st, cfg, 0.1 - t)
            n += 1
        out[f"t{k}_steps"] = np.array([n])
        out[f"t{k}_time"] = np.array([t])
        for fid in (0, 1, 4, 5, 8):
            out[f"t{k}_f{fid}"] = np.ascontiguousarray(mesh.interior(st.field(fid)).ravel())
    save("ensemble_ref.npz", **out)


GENERATORS = {"recon": g_recon, "riemann": g_riemann, "sod": g_sod, "vortex": g_vortex, "ghost": g_ghost,
              "jacobi": g_jacobi, "semi": g_semi, "liquid_gas": g_liquid_gas, "two_phase": g_two_phase,
              "igr": g_igr, "ensemble": g_ensemble, "body_force": g_body_force, "viscous": g_viscous,
              "capillary": g_capillary}

if __name__ == "__main__":
    assert po.available("ref"), "build oracle/_ref first (make -C oracle)"
    for name in sys.argv[1:] or GENERATORS:
        GENERATORS[name]()
