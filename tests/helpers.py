"""Shared builders for parity tests: random smooth/discontinuous states on
small meshes with every boundary rule, and relative error norms."""
from __future__ import annotations

import numpy as np

from semiimplicitfv_amd import (BoundaryCondition, IdealGasEOS, RectilinearMesh, SimulationConfig, SolutionState,
                                StiffenedGasEOS)
from semiimplicitfv_amd.cases import _fill
from tests.golden.problems import make_two_phase_case  # noqa: F401  (re-exported for the GPU tests)


def rel_errs(a: np.ndarray, b: np.ndarray):
    """(relative L1, relative Linf) of a against b."""
    d = np.abs(a - b)
    sb = np.abs(b).sum()
    mb = np.abs(b).max()
    l1 = d.sum() / sb if sb > 0 else d.sum()
    li = d.max() / mb if mb > 0 else d.max()
    return float(l1), float(li)


def make_case(dim, n, bcs=None, periods=(0, 0, 0), recon=2, rk=3, seed=0, nonuniform=False, eos_kind="ideal",
              ng=None, cfl=0.5, shock=True):
    """A random but well-posed state: smooth background + optional jump."""
    rng = np.random.default_rng(seed)
    cfg = SimulationConfig(dim=dim, nGhost=ng if ng else 3, RKOrder=rk, reconOrder=recon)
    cfg.explicitParams.cfl = cfl
    cfg.explicitParams.maxDt = 1.0
    cfg.validate()
    nodes = []
    for d in range(3):
        m = n[d] if d < dim else 1
        if d >= dim:
            nodes.append(np.array([0.0, 1.0]))
            continue
        if nonuniform:
            w = 1.0 + 0.5 * rng.random(m)
            x = np.concatenate([[0.0], np.cumsum(w)])
            x = x / x[-1]
        else:
            x = np.linspace(0.0, 1.0, m + 1)
        nodes.append(x)
    mesh = RectilinearMesh(cfg, nodes[0], nodes[1], nodes[2], periods=periods)
    if bcs is not None:
        for f, b in enumerate(bcs):
            mesh.setBoundaryCondition(f, b)
    st = SolutionState()
    st.allocate(mesh.totalCells(), cfg)
    if eos_kind == "stiff":
        eos = StiffenedGasEOS(4.4, 6e3, 1.0, cfg)
    else:
        eos = IdealGasEOS(1.4, 1.0, cfg)
    xc, yc, zc = mesh.centroids()
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * X + 0.3) * np.cos(2 * np.pi * Y) + 0.1 * np.cos(2 * np.pi * Z)
    p = 1.0 + 0.3 * np.cos(2 * np.pi * (X + Y + Z))
    if shock:
        jump = (X + 0.3 * Y + 0.2 * Z) < 0.45
        rho = np.where(jump, rho * 1.8, rho)
        p = np.where(jump, p * 3.0, p)
    u = [0.3 * np.sin(2 * np.pi * (Y + 0.1 * d)) + 0.1 * (d + 1) for d in range(dim)]
    _fill(mesh, st, eos, rho, u, p)
    return cfg, mesh, st, eos


ALL_BCS = [BoundaryCondition.Outflow, BoundaryCondition.Symmetry, BoundaryCondition.SlipWall,
           BoundaryCondition.NoSlipWall, BoundaryCondition.Periodic]


def make_semi_case(dim, n, bcs=None, periods=(0, 0, 0), recon=2, rk=3, seed=0, cfl=0.5, tol=1e-8, maxit=400,
                   nonuniform=False):
    """Single-phase ideal-gas state of make_case configured for the
    semi-implicit solver."""
    cfg, mesh, st, eos = make_case(dim, n, bcs=bcs, periods=periods, recon=recon, rk=rk, seed=seed,
                                   nonuniform=nonuniform)
    cfg.semiImplicit = True
    cfg.semiImplicitParams.cfl = cfl
    cfg.semiImplicitParams.maxDt = 1e-2
    cfg.semiImplicitParams.pressureTol = tol
    cfg.semiImplicitParams.maxPressureIters = maxit
    cfg.validate()
    return cfg, mesh, st, eos


def noise_floor_bar(ref, oracle, base=1e-12, factor=3.0):
    """(L1 bar, Linf bar) for one field of an ill-conditioned water/air case:
    the larger of `base` and `factor` times the discrepancy between the
    compiled reference and the oracle on the same inputs.  Those two are
    FP64 builds of the same algorithm that differ only in rounding (FMA
    contraction), so their distance is the rounding floor of the case."""
    l1, li = rel_errs(ref, oracle)
    return max(base, factor * l1), max(base, factor * li)
