"""Shared builders for parity tests: random smooth/discontinuous states on
small meshes with every boundary rule, and relative error norms."""
from __future__ import annotations

import numpy as np

from semiimplicitfv_amd import (BoundaryCondition, IdealGasEOS, RectilinearMesh, SimulationConfig, SolutionState,
                                StiffenedGasEOS)
from semiimplicitfv_amd.cases import _fill
from tests.golden.problems import make_case, make_two_phase_case  # noqa: F401  (re-exported for the tests)


def rel_errs(a: np.ndarray, b: np.ndarray):
    """(relative L1, relative Linf) of a against b."""
    d = np.abs(a - b)
    sb = np.abs(b).sum()
    mb = np.abs(b).max()
    l1 = d.sum() / sb if sb > 0 else d.sum()
    li = d.max() / mb if mb > 0 else d.max()
    return float(l1), float(li)


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
