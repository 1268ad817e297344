"""SemiImplicitFV hot path on MI355X (gfx950).

The compute path is libsifv_hip.so (HIP kernels + C ABI, include/sifv_gpu.h);
this package is the host-side mirror of the reference's interface for that
path: SimulationConfig, RectilinearMesh, SolutionState, the EOS and Riemann
solver objects, and GPU-backed ExplicitSolver / SemiImplicitSolver /
JacobiPressureSolver.
"""
from .config import (BoundaryCondition, ExplicitParams, IGRParams, MultiPhaseParams, PhaseEOS,
                     ReconstructionOrder, RiemannSolverType, SemiImplicitParams, SimulationConfig, VarSet)
from .mesh import RectilinearMesh, runtime_linspace, uniform_mesh
from .state import (HLLCSolver, IdealGasEOS, LFSolver, RusanovSolver, SolutionState, StiffenedGasEOS)
from .solvers import ExplicitSolver, JacobiPressureSolver, MGPCGPressureSolver, PCGPressureSolver, SemiImplicitSolver
from . import _abi
from .ensemble import RiemannEnsemble

__all__ = [
    "BoundaryCondition", "ExplicitParams", "IGRParams", "MultiPhaseParams", "PhaseEOS", "ReconstructionOrder",
    "RiemannSolverType", "SemiImplicitParams", "SimulationConfig", "VarSet", "RectilinearMesh", "runtime_linspace",
    "uniform_mesh", "HLLCSolver", "IdealGasEOS", "LFSolver", "RusanovSolver", "SolutionState", "StiffenedGasEOS",
    "ExplicitSolver", "JacobiPressureSolver", "MGPCGPressureSolver", "PCGPressureSolver", "SemiImplicitSolver", "RiemannEnsemble",
]
