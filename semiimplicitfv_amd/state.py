"""SolutionState mirror (include/SolutionState.hpp:25-129) plus the EOS and
Riemann-solver objects the reference's solver constructors take.

Fields are padded SoA float64 arrays indexed like RectilinearMesh::index.
"""
from __future__ import annotations

import numpy as np

from . import _abi
from .config import RiemannSolverType, SimulationConfig


class SolutionState:
    def __init__(self):
        self.dim_ = 3
        self.rho = self.rhoU = self.rhoV = self.rhoW = self.rhoE = None
        self.velU = self.velV = self.velW = self.pres = self.sigma = None
        self.alphaRho = []
        self.alpha = []

    def allocate(self, totalCells: int, config: SimulationConfig):
        """src/SolutionState.cpp:7-77 (backup/star arrays live inside the solvers)."""
        self.dim_ = config.dim
        z = lambda: np.zeros(totalCells, dtype=np.float64)
        self.rho, self.rhoU, self.rhoE = z(), z(), z()
        self.velU, self.pres, self.sigma = z(), z(), z()
        self.rhoV = z() if self.dim_ >= 2 else None
        self.velV = z() if self.dim_ >= 2 else None
        self.rhoW = z() if self.dim_ >= 3 else None
        self.velW = z() if self.dim_ >= 3 else None
        nph = config.multiPhaseParams.nPhases
        self.alphaRho = [z() for _ in range(nph)]
        self.alpha = [z() for _ in range(nph)]

    def dim(self):
        return self.dim_

    def size(self):
        return 0 if self.rho is None else len(self.rho)

    def field(self, fid: int):
        m = {_abi.F_RHO: self.rho, _abi.F_RHOU: self.rhoU, _abi.F_RHOV: self.rhoV, _abi.F_RHOW: self.rhoW,
             _abi.F_RHOE: self.rhoE, _abi.F_VELU: self.velU, _abi.F_VELV: self.velV, _abi.F_VELW: self.velW,
             _abi.F_PRES: self.pres, _abi.F_SIGMA: self.sigma}
        if fid in m:
            return m[fid]
        if _abi.F_ALPHARHO <= fid < _abi.F_ALPHARHO + _abi.MAX_PHASES:
            p = fid - _abi.F_ALPHARHO
            return self.alphaRho[p] if p < len(self.alphaRho) else None
        if _abi.F_ALPHA <= fid < _abi.F_ALPHA + _abi.MAX_PHASES:
            p = fid - _abi.F_ALPHA
            return self.alpha[p] if p < len(self.alpha) else None
        return None

    def field_ids(self):
        ids = [_abi.F_RHO, _abi.F_RHOU, _abi.F_RHOE, _abi.F_VELU, _abi.F_PRES, _abi.F_SIGMA]
        if self.dim_ >= 2:
            ids += [_abi.F_RHOV, _abi.F_VELV]
        if self.dim_ >= 3:
            ids += [_abi.F_RHOW, _abi.F_VELW]
        ids += [_abi.F_ALPHARHO + p for p in range(len(self.alphaRho))]
        ids += [_abi.F_ALPHA + p for p in range(len(self.alpha))]
        return sorted(ids)

    def copy(self) -> "SolutionState":
        s = SolutionState()
        s.dim_ = self.dim_
        for name in ("rho", "rhoU", "rhoV", "rhoW", "rhoE", "velU", "velV", "velW", "pres", "sigma"):
            a = getattr(self, name)
            setattr(s, name, None if a is None else a.copy())
        s.alphaRho = [a.copy() for a in self.alphaRho]
        s.alpha = [a.copy() for a in self.alpha]
        return s


class IdealGasEOS:
    """src/IdealGasEOS.cpp (vectorised over numpy arrays)."""

    kind = _abi.EOS_IDEAL

    def __init__(self, gamma: float, R: float = 287.0, config: SimulationConfig | None = None):
        self.gamma_ = gamma
        self.R_ = R
        self.dim_ = config.dim if config else 3

    def gamma(self):
        return self.gamma_

    def pInf(self):
        return 0.0

    def name(self):
        return "IdealGas"

    def toConservative(self, rho, u, p):
        """src/IdealGasEOS.cpp:66-82; u is a list of velocity arrays (len dim)."""
        mom = [rho * ui for ui in u]
        ke = 0.0
        for ui in u:
            ke = ke + ui * ui
        ke = ke * (0.5 * rho)
        E = p / (self.gamma_ - 1.0) + ke
        return rho, mom, E

    def to_c(self):
        e = _abi.Eos_C()
        e.kind, e.gamma, e.pInf = self.kind, self.gamma_, 0.0
        return e


class StiffenedGasEOS(IdealGasEOS):
    """src/StiffenedGasEOS.cpp"""

    kind = _abi.EOS_STIFFENED

    def __init__(self, gamma: float, pInf: float, R: float = 287.0, config: SimulationConfig | None = None):
        super().__init__(gamma, R, config)
        self.pInf_ = pInf

    def pInf(self):
        return self.pInf_

    def name(self):
        return "StiffenedGas"

    def toConservative(self, rho, u, p):
        """src/StiffenedGasEOS.cpp:67-84"""
        mom = [rho * ui for ui in u]
        ke = 0.0
        for ui in u:
            ke = ke + ui * ui
        ke = ke * (0.5 * rho)
        E = (p + self.gamma_ * self.pInf_) / (self.gamma_ - 1.0) + ke
        return rho, mom, E

    def to_c(self):
        e = _abi.Eos_C()
        e.kind, e.gamma, e.pInf = self.kind, self.gamma_, self.pInf_
        return e


class RiemannSolver:
    """Base of the named solver objects; the solvers map name() to the enum
    exactly as src/ExplicitSolver.cpp:31-34 does."""

    def __init__(self, eos, config: SimulationConfig | None = None):
        self.eos_ = eos

    def eos(self):
        return self.eos_

    def kind(self) -> RiemannSolverType:
        n = self.name()
        if n == "HLLC":
            return RiemannSolverType.HLLC
        if n == "Rusanov":
            return RiemannSolverType.Rusanov
        return RiemannSolverType.LF


class HLLCSolver(RiemannSolver):
    def name(self):
        return "HLLC"


class RusanovSolver(RiemannSolver):
    def name(self):
        return "Rusanov"


class LFSolver(RiemannSolver):
    def name(self):
        return "LF"
