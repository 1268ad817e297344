"""GPU-backed mirrors of the reference's solver classes.

ExplicitSolver / SemiImplicitSolver keep the reference constructor and
step() signatures (include/ExplicitSolver.hpp:23-46,
include/SemiImplicitSolver.hpp:24-46) and drive libsifv_hip.so through the
C ABI.  step() is a drop-in: it uploads the host SolutionState, advances it
on the GPU and writes every field back, exactly like the reference mutates
its SolutionState.  advance() keeps the state resident on the device for
many steps (the time-loop form used for measurements).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .config import SimulationConfig
from .mesh import RectilinearMesh
from .state import SolutionState


class _GpuSolverBase:
    def __init__(self, mesh: RectilinearMesh, riemannSolver, eos, igrSolver, config: SimulationConfig,
                 precision: str = "fp64"):
        config.validate()
        self._lib = _abi.load_library()
        self.mesh = mesh
        self.config = config
        self.eos = eos
        self.riemann = int(riemannSolver.kind())
        self.precision = {"fp64": _abi.FP64, "mixed": _abi.MIXED}[precision]
        h = C.c_void_p()
        mc = mesh.to_c()
        cc = config.to_c()
        ec = eos.to_c()
        _abi.check(self._lib.sifv_create(C.byref(mc), C.byref(cc), self.riemann, C.byref(ec), self.precision,
                                         C.byref(h)))
        self._h = h
        n = C.c_size_t()
        _abi.check(self._lib.sifv_total_cells(self._h, C.byref(n)))
        self.n = n.value
        if self.n != mesh.totalCells():
            raise RuntimeError("device and host disagree on totalCells")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sifv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state transfer
    def upload(self, state: SolutionState):
        for fid in state.field_ids():
            a = state.field(fid)
            a = np.ascontiguousarray(a, dtype=np.float64)
            _abi.check(self._lib.sifv_upload(self._h, fid, _abi.dptr(a), a.size))

    def download(self, state: SolutionState):
        for fid in state.field_ids():
            a = state.field(fid)
            buf = np.empty(self.n, dtype=np.float64)
            _abi.check(self._lib.sifv_download(self._h, fid, _abi.dptr(buf), buf.size))
            a[:] = buf

    def download_field(self, fid: int) -> np.ndarray:
        buf = np.empty(self.n, dtype=np.float64)
        _abi.check(self._lib.sifv_download(self._h, fid, _abi.dptr(buf), buf.size))
        return buf

    def upload_field(self, fid: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _abi.check(self._lib.sifv_upload(self._h, fid, _abi.dptr(a), a.size))

    # ---- reference-style helpers
    def fill_ghosts(self, varset: int):
        _abi.check(self._lib.sifv_fill_ghosts(self._h, varset))

    def fill_scalar_ghosts(self, fid: int):
        _abi.check(self._lib.sifv_fill_scalar_ghosts(self._h, fid))

    def acoustic_dt(self, cfl: float, maxDt: float) -> float:
        d = C.c_double()
        _abi.check(self._lib.sifv_acoustic_dt(self._h, cfl, maxDt, C.byref(d)))
        return d.value

    def advective_dt(self, cfl: float, maxDt: float) -> float:
        d = C.c_double()
        _abi.check(self._lib.sifv_advective_dt(self._h, cfl, maxDt, C.byref(d)))
        return d.value

    def cons_to_prim(self):
        _abi.check(self._lib.sifv_cons_to_prim(self._h))

    def synchronize(self):
        _abi.check(self._lib.sifv_synchronize(self._h))

    def set_tile(self, rows: int):
        _abi.check(self._lib.sifv_set_tile(self._h, rows))

    def set_stage_slabs(self, slabs: int):
        """3D: slabs of z segments a stage is pipelined over (0: built-in)."""
        _abi.check(self._lib.sifv_set_stage_slabs(self._h, slabs))

    def set_graphs(self, on: bool = True):
        """advance() replays each RK step from a captured HIP graph (on=True)
        or enqueues every launch (on=False, the default)."""
        _abi.check(self._lib.sifv_set_graphs(self._h, int(on)))

    def graph_replays(self) -> int:
        n = C.c_long()
        _abi.check(self._lib.sifv_graph_replays(self._h, C.byref(n)))
        return n.value

    def timing(self, on: bool = True):
        _abi.check(self._lib.sifv_timing_enable(self._h, int(on)))

    def read_timing(self, reset: bool = True):
        ms = C.c_double()
        n = C.c_long()
        _abi.check(self._lib.sifv_timing_read(self._h, C.byref(ms), C.byref(n), int(reset)))
        return ms.value, n.value


class ExplicitSolver(_GpuSolverBase):
    """ExplicitSolver (include/ExplicitSolver.hpp:21-82) on the GPU."""

    def step(self, config: SimulationConfig, mesh: RectilinearMesh, state: SolutionState,
             targetDt: float = -1.0) -> float:
        self.upload(state)
        dt = self.step_resident(config, targetDt)
        self.download(state)
        return dt

    def step_resident(self, config: SimulationConfig, targetDt: float = -1.0) -> float:
        """One step on the device-resident state (no host transfer)."""
        d = C.c_double()
        cc = config.to_c()
        _abi.check(self._lib.sifv_explicit_step(self._h, C.byref(cc), float(targetDt), C.byref(d)))
        return d.value

    def advance(self, config: SimulationConfig, nsteps: int, endTime: float, time: float = 0.0):
        """Up to nsteps steps of runTimeLoop (targetDt = endTime - time, no
        output in between) on the resident state; returns (time, steps)."""
        t = C.c_double(time)
        s = C.c_int()
        cc = config.to_c()
        _abi.check(self._lib.sifv_explicit_advance(self._h, C.byref(cc), int(nsteps), float(endTime), C.byref(t),
                                                   C.byref(s)))
        return t.value, s.value


    # ---- discrete adjoint (no reference counterpart) ----------------------
    def adjoint_forward(self, config: SimulationConfig, nsteps: int, checkpoint_every: int = 20):
        """Runs nsteps fixed-dt steps from the resident state, keeping
        checkpoints for adjoint_backward."""
        cc = config.to_c()
        _abi.check(self._lib.sifv_adjoint_forward(self._h, C.byref(cc), int(nsteps), int(checkpoint_every)))

    def adjoint_backward(self, seed):
        """seed: 4 padded arrays (rho, rhoU, rhoV, rhoE), the adjoint of the
        final state.  -> (4 padded gradient arrays w.r.t. the state at the
        start of adjoint_forward, d/dgamma)."""
        n = self.mesh.totalCells()
        s = np.ascontiguousarray(np.concatenate([np.asarray(a, dtype=np.float64).ravel() for a in seed]))
        if s.size != 4 * n:
            raise ValueError("seed must hold 4 fields of totalCells values")
        g = np.empty(4 * n)
        dg = C.c_double()
        _abi.check(self._lib.sifv_adjoint_backward_host(self._h, _abi.dptr(s), _abi.dptr(g), C.byref(dg)))
        return [g[q * n:(q + 1) * n] for q in range(4)], dg.value

    def state_device_ptr(self) -> int:
        p = C.c_void_p()
        _abi.check(self._lib.sifv_state_device_ptr(self._h, C.byref(p)))
        return p.value


class SemiImplicitSolver(_GpuSolverBase):
    """SemiImplicitSolver (include/SemiImplicitSolver.hpp:22-96) on the GPU,
    with the Jacobi, the PCG or the multigrid-preconditioned PCG pressure
    solver (chosen by the pressure solver's name(), as the reference
    dispatches on the object passed in)."""

    _PS = {"Jacobi": 0, "PCG": 1, "MGPCG": 2}

    def __init__(self, mesh, riemannSolver, pressureSolver, eos, igrSolver, config, precision="fp64"):
        super().__init__(mesh, riemannSolver, eos, igrSolver, config, precision)
        self.pressureSolver_ = pressureSolver
        self.lastIters_ = 0
        name = pressureSolver.name() if pressureSolver is not None else "Jacobi"
        if name not in self._PS:
            self.close()
            raise ValueError(f"no GPU pressure solver named {name!r}")
        if self._PS[name]:
            try:
                _abi.check(self._lib.sifv_set_pressure_solver(self._h, self._PS[name]))
            except Exception:
                self.close()
                raise

    def step(self, config, mesh, state, targetDt=-1.0):
        self.upload(state)
        dt = self.step_resident(config, targetDt)
        self.download(state)
        return dt

    def step_resident(self, config, targetDt=-1.0):
        d = C.c_double()
        it = C.c_int()
        cc = config.to_c()
        _abi.check(self._lib.sifv_semi_implicit_step(self._h, C.byref(cc), float(targetDt), C.byref(d),
                                                     C.byref(it)))
        self.lastIters_ = it.value
        return d.value

    def lastPressureIterations(self):
        return self.lastIters_


class JacobiPressureSolver:
    """PressureSolver::solve (include/PressureSolver.hpp:17-40) backed by the
    GPU Jacobi sweep.  Needs a handle for the mesh (any solver built on the
    same mesh)."""

    def __init__(self, gpu_solver: _GpuSolverBase | None = None):
        self.gpu = gpu_solver

    def name(self):
        return "Jacobi"

    def solve(self, mesh, rho, rhoc2, rhs, pressure, dt, tolerance, maxIter, history=None):
        lib = _abi.load_library()
        it = C.c_int()
        hist = history if history is not None else None
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, rhoc2, rhs)]
        if not (pressure.flags.c_contiguous and pressure.dtype == np.float64):
            raise ValueError("pressure must be a contiguous float64 array (updated in place)")
        _abi.check(lib.sifv_jacobi_solve(self.gpu._h, _abi.dptr(arrs[0]), _abi.dptr(arrs[1]), _abi.dptr(arrs[2]),
                                         _abi.dptr(pressure), float(dt), float(tolerance), int(maxIter),
                                         C.byref(it), _abi.dptr(hist) if hist is not None else None))
        return it.value


class PCGPressureSolver:
    """PressureSolver::solve (include/PressureSolver.hpp:17-40) in the role of
    HyprePressureSolver (src/HyprePressureSolver.cpp:312-408): conjugate
    gradients on the GPU (sifv_pcg_solve), stopping on the relative two-norm
    of the residual.  Needs a handle for the mesh (any solver built on the same
    mesh) for standalone solves; passed to SemiImplicitSolver it selects the
    PCG solve of the semi-implicit step."""

    def __init__(self, gpu_solver: _GpuSolverBase | None = None):
        self.gpu = gpu_solver

    def name(self):
        return "PCG"

    def solve(self, mesh, rho, rhoc2, rhs, pressure, dt, tolerance, maxIter):
        lib = _abi.load_library()
        it = C.c_int()
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, rhoc2, rhs)]
        if not (pressure.flags.c_contiguous and pressure.dtype == np.float64):
            raise ValueError("pressure must be a contiguous float64 array (updated in place)")
        _abi.check(lib.sifv_pcg_solve(self.gpu._h, _abi.dptr(arrs[0]), _abi.dptr(arrs[1]), _abi.dptr(arrs[2]),
                                      _abi.dptr(pressure), float(dt), float(tolerance), int(maxIter), C.byref(it)))
        return it.value


class MGPCGPressureSolver(PCGPressureSolver):
    """PressureSolver::solve (include/PressureSolver.hpp:17-40) in the role of
    HyprePressureSolver's StructPCG + PFMG pairing (src/HyprePressureSolver.cpp:
    123-138, 312-408): conjugate gradients on the GPU preconditioned by one
    geometric-multigrid V-cycle (sifv_mgpcg_solve), stopping on the relative
    two-norm of the residual.  Passed to SemiImplicitSolver it selects the
    multigrid PCG solve of the semi-implicit step."""

    def name(self):
        return "MGPCG"

    def solve(self, mesh, rho, rhoc2, rhs, pressure, dt, tolerance, maxIter):
        lib = _abi.load_library()
        it = C.c_int()
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, rhoc2, rhs)]
        if not (pressure.flags.c_contiguous and pressure.dtype == np.float64):
            raise ValueError("pressure must be a contiguous float64 array (updated in place)")
        _abi.check(lib.sifv_mgpcg_solve(self.gpu._h, _abi.dptr(arrs[0]), _abi.dptr(arrs[1]), _abi.dptr(arrs[2]),
                                        _abi.dptr(pressure), float(dt), float(tolerance), int(maxIter),
                                        C.byref(it)))
        return it.value
