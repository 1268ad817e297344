"""Batched 1D runs on the GPU (sifv_ensemble_*, include/sifv_gpu.h).

No counterpart in the reference, which runs one problem per process: each
tube here is what ExplicitSolver + runTimeLoop would do for one 1D shock
tube (outflow ends, WENO5, SSP-RK3, no output), and all tubes advance to the
end time in one persistent launch.  Fields are packed tube after tube
(interior cells only).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .cases import riemann_by_name, shock_tube


class RiemannEnsemble:
    def __init__(self, tubes, config, riemannSolver, eos, precision: str = "fp64"):
        """tubes: list of (mesh, state) pairs of 1D runs sharing config / EOS."""
        self._lib = _abi.load_library()
        self.meshes = [m for m, _ in tubes]
        self.ncells = np.array([m.nx() for m in self.meshes], dtype=np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(self.ncells)]).astype(np.int64)
        nodes = np.concatenate([m.xNodes for m in self.meshes]).astype(np.float64)
        cc = config.to_c()
        ec = eos.to_c()
        h = C.c_void_p()
        prec = {"fp64": _abi.FP64, "mixed": _abi.MIXED}[precision]
        _abi.check(self._lib.sifv_ensemble_create(len(tubes), self.ncells.ctypes.data_as(C.POINTER(C.c_int)),
                                                  _abi.dptr(nodes), C.byref(cc), riemannSolver.kind(),
                                                  C.byref(ec), prec, C.byref(h)))
        self._h = h
        self.ntot = int(self.offsets[-1])
        self.upload([s for _, s in tubes])

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sifv_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _pack(self, states, fid):
        return np.ascontiguousarray(
            np.concatenate([m.interior(s.field(fid)).ravel() for m, s in zip(self.meshes, states)]), dtype=np.float64)

    def upload(self, states):
        for fid in (_abi.F_RHO, _abi.F_RHOU, _abi.F_RHOE, _abi.F_VELU, _abi.F_PRES):
            a = self._pack(states, fid)
            _abi.check(self._lib.sifv_ensemble_upload(self._h, fid, _abi.dptr(a), a.size))

    def download(self, fid: int):
        """Per-tube interior arrays of one field."""
        a = np.empty(self.ntot)
        _abi.check(self._lib.sifv_ensemble_download(self._h, fid, _abi.dptr(a), a.size))
        return [a[self.offsets[t]:self.offsets[t + 1]] for t in range(len(self.ncells))]

    def set_shape(self, block: int, wg_per_cu: int = 4):
        _abi.check(self._lib.sifv_ensemble_set_shape(self._h, int(block), int(wg_per_cu)))

    def run(self, endTime: float, maxSteps: int = 10 ** 7):
        """-> (steps per tube, final time per tube, kernel ms)."""
        n = len(self.ncells)
        steps = np.zeros(n, dtype=np.int32)
        times = np.zeros(n)
        ms = C.c_double()
        _abi.check(self._lib.sifv_ensemble_run(self._h, float(endTime), int(maxSteps),
                                               steps.ctypes.data_as(C.POINTER(C.c_int)), _abi.dptr(times),
                                               C.byref(ms)))
        return steps, times, ms.value

    def autotune(self, endTime: float, maxSteps: int):
        """Times every tile shape on the current state (restored afterwards)
        and keeps the fastest -> (block, wg_per_cu, ms)."""
        b, w, ms = C.c_int(), C.c_int(), C.c_double()
        _abi.check(self._lib.sifv_ensemble_autotune(self._h, float(endTime), int(maxSteps), C.byref(b), C.byref(w),
                                                    C.byref(ms)))
        return b.value, w.value, ms.value


def build_tubes(params, indices=None, riemann: str = "hllc", cfl: float = 0.6, gamma: float = 1.4):
    """(config, eos, Riemann solver, [(mesh, state)]) of rows of a
    riemann_ensemble_params table."""
    idx = range(len(params["ncells"])) if indices is None else indices
    tubes = []
    cfg = eos = rs = None
    for t in idx:
        cfg, mesh, st, eos, rs = shock_tube(int(params["ncells"][t]),
                                            (params["rho"][t, 0], params["u"][t, 0], params["p"][t, 0]),
                                            (params["rho"][t, 1], params["u"][t, 1], params["p"][t, 1]),
                                            x0=float(params["x0"][t]), riemann=riemann, cfl=cfl, gamma=gamma)
        tubes.append((mesh, st))
    return cfg, eos, rs, tubes
