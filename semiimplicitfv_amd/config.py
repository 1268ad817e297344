"""Python mirror of SimulationConfig (include/SimulationConfig.hpp:13-210).

Field names, defaults, requiredGhostCells() and validate() follow the
reference; validate() raises ValueError where the reference throws
std::invalid_argument.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass, field
from typing import List

from . import _abi


class ReconstructionOrder(enum.IntEnum):
    """include/SimulationConfig.hpp:13-20"""
    WENO1 = 0
    WENO3 = 1
    WENO5 = 2
    UPWIND1 = 3
    UPWIND3 = 4
    UPWIND5 = 5


class RiemannSolverType(enum.IntEnum):
    """include/RiemannSolver.hpp:16"""
    LF = 0
    Rusanov = 1
    HLLC = 2


class BoundaryCondition(enum.IntEnum):
    """include/RectilinearMesh.hpp:14-20"""
    Symmetry = 0
    Outflow = 1
    Periodic = 2
    SlipWall = 3
    NoSlipWall = 4


class VarSet(enum.IntEnum):
    """include/SolutionState.hpp:15-19"""
    CONS = 0
    PRIM = 1
    SIGMA = 2


@dataclass
class ExplicitParams:
    cfl: float = 0.8
    constDt: float = -1.0
    maxDt: float = 1e-2
    minDt: float = 1e-12


@dataclass
class SemiImplicitParams:
    cfl: float = 0.8
    maxDt: float = 1e-2
    minDt: float = 1e-12
    maxPressureIters: int = 100
    pressureTol: float = 1e-8


@dataclass
class IGRParams:
    alphaCoeff: float = 1.0
    IGRIters: int = 5
    IGRWarmStartIters: int = 50


@dataclass
class PhaseEOS:
    gamma: float = 1.4
    pInf: float = 0.0


@dataclass
class MultiPhaseParams:
    nPhases: int = 0
    phases: List[PhaseEOS] = field(default_factory=list)
    alphaMin: float = 1e-8


@dataclass
class BodyForceParams:
    a: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    b: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    c: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    d: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])


@dataclass
class ViscousParams:
    mu: float = 0.0
    phaseMu: List[float] = field(default_factory=list)


@dataclass
class SurfaceTensionParams:
    sigma: float = 0.0
    epsGradAlpha: float = 1e-8


@dataclass
class SimulationConfig:
    dim: int = 3
    nGhost: int = 2
    RKOrder: int = 1
    useIGR: bool = False
    semiImplicit: bool = False
    reconOrder: ReconstructionOrder = ReconstructionOrder.WENO1
    wenoEps: float = 1e-6
    step: int = 0
    time: float = 0.0
    explicitParams: ExplicitParams = field(default_factory=ExplicitParams)
    semiImplicitParams: SemiImplicitParams = field(default_factory=SemiImplicitParams)
    igrParams: IGRParams = field(default_factory=IGRParams)
    multiPhaseParams: MultiPhaseParams = field(default_factory=MultiPhaseParams)
    bodyForceParams: BodyForceParams = field(default_factory=BodyForceParams)
    viscousParams: ViscousParams = field(default_factory=ViscousParams)
    surfaceTensionParams: SurfaceTensionParams = field(default_factory=SurfaceTensionParams)

    def isMultiPhase(self) -> bool:
        return self.multiPhaseParams.nPhases > 0

    def hasViscosity(self) -> bool:
        return self.viscousParams.mu > 0.0 or len(self.viscousParams.phaseMu) > 0

    def hasSurfaceTension(self) -> bool:
        return self.surfaceTensionParams.sigma > 0.0

    def hasBodyForce(self) -> bool:
        bf = self.bodyForceParams
        return any(bf.a[i] != 0.0 or bf.b[i] != 0.0 for i in range(3))

    def requiredGhostCells(self) -> int:
        """include/SimulationConfig.hpp:108-121"""
        o = ReconstructionOrder(self.reconOrder)
        if o in (ReconstructionOrder.WENO1, ReconstructionOrder.UPWIND1):
            return 1
        if o in (ReconstructionOrder.WENO3, ReconstructionOrder.UPWIND3):
            return 2
        return 3

    def validate(self) -> None:
        """include/SimulationConfig.hpp:123-209"""
        if self.dim < 1 or self.dim > 3:
            raise ValueError(f"dim must be 1, 2, or 3 (got {self.dim})")
        if self.RKOrder < 1 or self.RKOrder > 3:
            raise ValueError(f"RKOrder must be 1, 2, or 3 (got {self.RKOrder})")
        req = self.requiredGhostCells()
        if self.nGhost < req:
            raise ValueError(f"nGhost={self.nGhost} is too small for chosen reconOrder (need >= {req})")
        if self.semiImplicit:
            p = self.semiImplicitParams
            if p.cfl <= 0:
                raise ValueError("semiImplicitParams.cfl must be > 0")
            if p.minDt <= 0:
                raise ValueError("semiImplicitParams.minDt must be > 0")
            if p.maxDt <= p.minDt:
                raise ValueError("semiImplicitParams.maxDt must be > minDt")
            if p.pressureTol <= 0:
                raise ValueError("semiImplicitParams.pressureTol must be > 0")
            if p.maxPressureIters <= 0:
                raise ValueError("semiImplicitParams.maxPressureIters must be > 0")
        else:
            p = self.explicitParams
            if p.cfl <= 0:
                raise ValueError("explicitParams.cfl must be > 0")
            if p.minDt <= 0:
                raise ValueError("explicitParams.minDt must be > 0")
            if p.maxDt <= p.minDt:
                raise ValueError("explicitParams.maxDt must be > minDt")
        if self.useIGR:
            p = self.igrParams
            if p.alphaCoeff <= 0:
                raise ValueError("igrParams.alphaCoeff must be > 0")
            if p.IGRIters <= 0:
                raise ValueError("igrParams.IGRIters must be > 0")
            if p.IGRWarmStartIters < 0:
                raise ValueError("igrParams.IGRWarmStartIters must be >= 0")
        if self.isMultiPhase():
            mp = self.multiPhaseParams
            if len(mp.phases) != mp.nPhases:
                raise ValueError("multiPhaseParams.phases.size() must equal nPhases")
            for ph, e in enumerate(mp.phases):
                if e.gamma <= 1.0:
                    raise ValueError(f"multiPhaseParams.phases[{ph}].gamma must be > 1")
                if e.pInf < 0.0:
                    raise ValueError(f"multiPhaseParams.phases[{ph}].pInf must be >= 0")
            if mp.alphaMin <= 0.0:
                raise ValueError("multiPhaseParams.alphaMin must be > 0")
        if self.viscousParams.mu < 0.0:
            raise ValueError("viscousParams.mu must be >= 0")
        if self.viscousParams.phaseMu:
            if not self.isMultiPhase():
                raise ValueError("viscousParams.phaseMu requires multi-phase (nPhases >= 2)")
            if len(self.viscousParams.phaseMu) != self.multiPhaseParams.nPhases:
                raise ValueError("viscousParams.phaseMu.size() must equal nPhases")
            for ph, mu in enumerate(self.viscousParams.phaseMu):
                if mu < 0.0:
                    raise ValueError(f"viscousParams.phaseMu[{ph}] must be >= 0")
        if self.surfaceTensionParams.sigma < 0.0:
            raise ValueError("surfaceTensionParams.sigma must be >= 0")
        if self.surfaceTensionParams.epsGradAlpha <= 0.0:
            raise ValueError("surfaceTensionParams.epsGradAlpha must be > 0")

    def to_c(self) -> _abi.Config_C:
        c = _abi.Config_C()
        c.dim = self.dim
        c.nGhost = self.nGhost
        c.RKOrder = self.RKOrder
        c.useIGR = int(bool(self.useIGR))
        c.semiImplicit = int(bool(self.semiImplicit))
        c.reconOrder = int(self.reconOrder)
        c.wenoEps = self.wenoEps
        c.step = self.step
        c.time = self.time
        e = self.explicitParams
        c.explicitParams.cfl, c.explicitParams.constDt = e.cfl, e.constDt
        c.explicitParams.maxDt, c.explicitParams.minDt = e.maxDt, e.minDt
        s = self.semiImplicitParams
        c.semiImplicitParams.cfl, c.semiImplicitParams.maxDt = s.cfl, s.maxDt
        c.semiImplicitParams.minDt = s.minDt
        c.semiImplicitParams.maxPressureIters = s.maxPressureIters
        c.semiImplicitParams.pressureTol = s.pressureTol
        g = self.igrParams
        c.igrParams.alphaCoeff, c.igrParams.IGRIters = g.alphaCoeff, g.IGRIters
        c.igrParams.IGRWarmStartIters = g.IGRWarmStartIters
        mp = self.multiPhaseParams
        c.multiPhaseParams.nPhases = mp.nPhases
        for i, ph in enumerate(mp.phases[:_abi.MAX_PHASES]):
            c.multiPhaseParams.phases[i].gamma = ph.gamma
            c.multiPhaseParams.phases[i].pInf = ph.pInf
        c.multiPhaseParams.alphaMin = mp.alphaMin
        bf = self.bodyForceParams
        for i in range(3):
            c.bodyForceParams.a[i], c.bodyForceParams.b[i] = bf.a[i], bf.b[i]
            c.bodyForceParams.c[i], c.bodyForceParams.d[i] = bf.c[i], bf.d[i]
        c.viscousMu = self.viscousParams.mu
        pm = list(self.viscousParams.phaseMu)
        if len(pm) > _abi.MAX_PHASES:
            raise ValueError("viscousParams.phaseMu: too many phases")
        c.viscousNPhaseMu = len(pm)
        for i, m in enumerate(pm):
            c.viscousPhaseMu[i] = m
        c.surfaceTensionSigma = self.surfaceTensionParams.sigma
        c.surfaceTensionEps = self.surfaceTensionParams.epsGradAlpha
        return c
