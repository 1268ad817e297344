"""VTK output of a GPU solver's resident state and the time loop that writes it.

VTKSession mirrors the reference's VTKSession on one rank
(src/VTKSession.cpp:11-63): a .pvd collection, one snapshot_<n>/ directory
per output holding <base>_<n>_r0.vtr and <base>_<n>.pvtr.  The files are
written by the library (sifv_write_vtr / _pvtr / _pvd, the reference's
VTKWriter format), which downloads the fields it needs at output time.

run_time_loop_vtk mirrors runTimeLoop (src/RKTimeStepping.cpp:241-356): a
snapshot at t = 0, dt clamped to the next output time and the end time,
a snapshot whenever an output time is reached, the NaN check at each output
and the final close.
"""
from __future__ import annotations

import os

import numpy as np

from . import _abi


class VTKSession:
    def __init__(self, solver, baseName: str, dir: str = "."):
        self.solver = solver
        self.baseName = baseName
        self.dir = dir
        self.fileNum = 0
        self.lib = _abi.load_library()
        _abi.check(self.lib.sifv_write_pvd(self._pvd(), b"w", 0.0, b""))

    def _pvd(self) -> bytes:
        return os.path.join(self.dir, self.baseName + ".pvd").encode()

    def write(self, time: float):
        """Snapshot of the solver's resident state at `time`."""
        snap = f"snapshot_{self.fileNum}"
        vtr = f"{self.baseName}_{self.fileNum}_r0.vtr"
        pvtr = f"{self.baseName}_{self.fileNum}.pvtr"
        h = self.solver._h
        _abi.check(self.lib.sifv_write_vtr(h, os.path.join(self.dir, snap, vtr).encode(), 0))
        _abi.check(self.lib.sifv_write_pvtr(h, os.path.join(self.dir, snap, pvtr).encode(), vtr.encode()))
        _abi.check(self.lib.sifv_write_pvd(self._pvd(), b"a", float(time), f"{snap}/{pvtr}".encode()))
        self.fileNum += 1

    def finalize(self):
        _abi.check(self.lib.sifv_write_pvd(self._pvd(), b"close", 0.0, b""))


def _has_nan(solver, dim: int) -> str | None:
    """The runTimeLoop NaN check (src/RKTimeStepping.cpp:281-291) on the
    interior cells of the resident state."""
    mesh = solver.mesh
    fields = [(_abi.F_RHO, "rho"), (_abi.F_RHOU, "rhoU"), (_abi.F_RHOE, "rhoE"), (_abi.F_PRES, "pres")]
    if dim >= 2:
        fields.append((_abi.F_RHOV, "rhoV"))
    if dim >= 3:
        fields.append((_abi.F_RHOW, "rhoW"))
    for fid, name in fields:
        if np.isnan(mesh.interior(solver.download_field(fid))).any():
            return name
    return None


def run_time_loop_vtk(config, solver, vtk: VTKSession, stepFn, endTime: float, outputInterval: float,
                      checkNaN: bool = True):
    """runTimeLoop with the VTK session writing the resident state; stepFn(targetDt)
    -> dt advances the solver (e.g. solver.step_resident).  Returns (time, steps)."""
    vtk.write(0.0)
    time = 0.0
    nextOutput = outputInterval
    steps = 0
    while time < endTime:
        targetDt = endTime - time
        if nextOutput < endTime:
            targetDt = min(targetDt, nextOutput - time)
        config.time = time
        dt = stepFn(targetDt)
        time += dt
        config.step += 1
        steps += 1
        if time >= nextOutput - 1e-12 * outputInterval:
            vtk.write(time)
            nextOutput += outputInterval
            if checkNaN:
                bad = _has_nan(solver, config.dim)
                if bad:
                    vtk.finalize()
                    raise FloatingPointError(f"NaN detected in field '{bad}' at step {config.step}, t = {time}")
    vtk.finalize()
    return time, steps
