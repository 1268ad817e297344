"""ctypes bindings of the CPU checkers.

TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline leg).  Two interchangeable implementations share one Python
interface:

  impl="c"   liboracle.so          - the C restatement (oracle/sifv_oracle.c)
  impl="ref" _ref/libsifv_ref.so   - the unmodified reference sources plus a
                                     thin C shim (oracle/ref_driver.cpp)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from semiimplicitfv_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsifv_ref.so")
REF_BENCH = os.path.join(HERE, "_ref", "ref_bench")

_libs = {}


def build(ref: bool = True, quiet: bool = True):
    """Build liboracle.so (always) and oracle/_ref (when the reference checkout
    is present).  Returns the make exit status."""
    tgt = "all" if ref else "oracle"
    r = subprocess.run(["make", "-C", HERE, "-j8", tgt], stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0 or not quiet:
        print(r.stdout.decode(errors="replace"))
    return r.returncode


def available(impl: str) -> bool:
    return os.path.exists(ORACLE_SO if impl == "c" else REF_SO)


def _lib(impl: str):
    if impl in _libs:
        return _libs[impl]
    path = ORACLE_SO if impl == "c" else REF_SO
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (run make -C oracle)")
    lib = C.CDLL(path)
    pre = "or_" if impl == "c" else "ref_"
    P = C.POINTER
    vp, d, i = C.c_void_p, C.c_double, C.c_int
    pd = P(d)
    ppd = P(pd)
    sig = {
        "create": (vp, [P(_abi.Mesh_C), P(_abi.Config_C), i, P(_abi.Eos_C)]),
        "destroy": (None, [vp]),
        "total_cells": (C.c_size_t, [vp]),
        "last_error": (C.c_char_p, []),
        "explicit_step": (d, [vp, ppd, P(_abi.Config_C), d]),
        "semi_step": (d, [vp, ppd, P(_abi.Config_C), d, P(i)]),
        "fill_ghosts": (None, [vp, ppd, i]),
        "fill_scalar": (None, [vp, pd]),
        "cons_to_prim": (None, [vp, ppd, P(_abi.Config_C)]),
        "acoustic_dt": (d, [vp, ppd, P(_abi.Config_C), d, d]),
        "advective_dt": (d, [vp, ppd, d, d]),
        "igr": (None, [vp, ppd, P(_abi.Config_C)]),
        "riemann_flux": (None, [i, pd, pd, i, i, i, i, i, pd]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, pre + name)
        f.restype = res
        f.argtypes = args
    if impl == "c":
        lib.or_jacobi.restype = i
        lib.or_jacobi.argtypes = [vp, pd, pd, pd, pd, d, d, i, pd]
        lib.or_pcg.restype = i
        lib.or_pcg.argtypes = [vp, pd, pd, pd, pd, d, d, i]
        lib.or_mgpcg.restype = i
        lib.or_mgpcg.argtypes = [vp, pd, pd, pd, pd, d, d, i]
        lib.or_mg_precond.restype = i
        lib.or_mg_precond.argtypes = [vp, pd, pd, d, pd, pd]
        lib.or_set_pressure_solver.restype = None
        lib.or_set_pressure_solver.argtypes = [vp, i]
        lib.or_recon.restype = d
        lib.or_recon.argtypes = [i, i, pd, d]
        lib.or_compute_rhs.restype = None
        lib.or_compute_rhs.argtypes = [vp, ppd, P(_abi.Config_C), ppd]
    else:
        lib.ref_jacobi.restype = i
        lib.ref_jacobi.argtypes = [vp, pd, pd, pd, pd, d, d, i]
        lib.ref_recon_line.restype = i
        lib.ref_recon_line.argtypes = [i, d, i, i, pd, pd, pd]
        lib.ref_write_vtr.restype = i
        lib.ref_write_vtr.argtypes = [vp, ppd, C.c_char_p, i]
        lib.ref_write_pvtr.restype = i
        lib.ref_write_pvtr.argtypes = [vp, C.c_char_p, C.c_char_p]
        lib.ref_write_pvd.restype = i
        lib.ref_write_pvd.argtypes = [C.c_char_p, C.c_char_p, d, C.c_char_p]
    _libs[impl] = (lib, pre)
    return _libs[impl]


def _fieldptrs(state):
    arr = (C.POINTER(C.c_double) * _abi.N_FIELD_SLOTS)()
    keep = []
    for fid in range(_abi.N_FIELD_SLOTS):
        a = state.field(fid)
        if a is not None:
            assert a.dtype == np.float64 and a.flags.c_contiguous
            arr[fid] = _abi.dptr(a)
            keep.append(a)
    return arr, keep


class CpuSolver:
    """One solver instance of the C restatement or the compiled reference."""

    def __init__(self, mesh, config, riemann: int, eos, impl: str = "c"):
        self.impl = impl
        self.lib, self.pre = _lib(impl)
        self._mc = mesh.to_c()
        self._cc = config.to_c()
        self._ec = eos.to_c()
        self.h = self._fn("create")(C.byref(self._mc), C.byref(self._cc), int(riemann), C.byref(self._ec))
        if not self.h:
            raise RuntimeError(f"{impl} create failed: {self._fn('last_error')().decode()}")
        self.n = self._fn("total_cells")(self.h)

    def _fn(self, name):
        return getattr(self.lib, self.pre + name)

    def close(self):
        if self.h:
            self._fn("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def explicit_step(self, state, config, targetDt=-1.0):
        f, keep = _fieldptrs(state)
        cc = config.to_c()
        return self._fn("explicit_step")(self.h, f, C.byref(cc), float(targetDt))

    def semi_step(self, state, config, targetDt=-1.0):
        f, keep = _fieldptrs(state)
        cc = config.to_c()
        it = C.c_int()
        dt = self._fn("semi_step")(self.h, f, C.byref(cc), float(targetDt), C.byref(it))
        return dt, it.value

    def fill_ghosts(self, state, varset):
        f, keep = _fieldptrs(state)
        self._fn("fill_ghosts")(self.h, f, int(varset))

    def fill_scalar(self, a):
        self._fn("fill_scalar")(self.h, _abi.dptr(a))

    def cons_to_prim(self, state, config):
        f, keep = _fieldptrs(state)
        cc = config.to_c()
        self._fn("cons_to_prim")(self.h, f, C.byref(cc))

    def acoustic_dt(self, state, config, cfl, maxDt):
        f, keep = _fieldptrs(state)
        cc = config.to_c()
        return self._fn("acoustic_dt")(self.h, f, C.byref(cc), float(cfl), float(maxDt))

    def advective_dt(self, state, cfl, maxDt):
        f, keep = _fieldptrs(state)
        return self._fn("advective_dt")(self.h, f, float(cfl), float(maxDt))

    def igr(self, state, config):
        f, keep = _fieldptrs(state)
        cc = config.to_c()
        self._fn("igr")(self.h, f, C.byref(cc))

    # ---- the compiled reference's VTK writer (VTKWriter, src/VTKWriter.cpp) ----
    def write_vtr(self, state, filename: str, rank: int = 0):
        assert self.impl == "ref"
        f, keep = _fieldptrs(state)
        if self.lib.ref_write_vtr(self.h, f, filename.encode(), int(rank)) != 0:
            raise RuntimeError(self.lib.ref_last_error().decode())

    def write_pvtr(self, filename: str, piece_file: str):
        assert self.impl == "ref"
        if self.lib.ref_write_pvtr(self.h, filename.encode(), piece_file.encode()) != 0:
            raise RuntimeError(self.lib.ref_last_error().decode())

    def write_pvd(self, filename: str, mode: str, time: float = 0.0, data_file: str = ""):
        assert self.impl == "ref"
        if self.lib.ref_write_pvd(filename.encode(), mode.encode(), float(time), data_file.encode()) != 0:
            raise RuntimeError(self.lib.ref_last_error().decode())

    def jacobi(self, rho, rhoc2, rhs, p, dt, tol, maxIter, hist=None):
        if self.impl == "c":
            return self.lib.or_jacobi(self.h, _abi.dptr(rho), _abi.dptr(rhoc2), _abi.dptr(rhs), _abi.dptr(p), dt,
                                      tol, maxIter, _abi.dptr(hist) if hist is not None else None)
        return self.lib.ref_jacobi(self.h, _abi.dptr(rho), _abi.dptr(rhoc2), _abi.dptr(rhs), _abi.dptr(p), dt, tol,
                                   maxIter)

    def pcg(self, rho, rhoc2, rhs, p, dt, tol, maxIter):
        """PCG pressure solve (oracle only): p updated in place (interior)."""
        assert self.impl == "c"
        return self.lib.or_pcg(self.h, _abi.dptr(rho), _abi.dptr(rhoc2), _abi.dptr(rhs), _abi.dptr(p), dt, tol,
                               maxIter)

    def mgpcg(self, rho, rhoc2, rhs, p, dt, tol, maxIter):
        """PCG with the multigrid V-cycle preconditioner (oracle only)."""
        assert self.impl == "c"
        return self.lib.or_mgpcg(self.h, _abi.dptr(rho), _abi.dptr(rhoc2), _abi.dptr(rhs), _abi.dptr(p), dt, tol,
                                 maxIter)

    def mg_precond(self, rho, rhoc2, dt, r, z):
        """z = one multigrid V-cycle applied to r (oracle only); returns the
        number of coarse levels, -1 when the mesh is refused."""
        assert self.impl == "c"
        return self.lib.or_mg_precond(self.h, _abi.dptr(rho), _abi.dptr(rhoc2), dt, _abi.dptr(r), _abi.dptr(z))

    def set_pressure_solver(self, kind: int):
        """Pressure solver of semi_step (oracle only): 0 Jacobi, 1 PCG, 2 multigrid PCG."""
        assert self.impl == "c"
        self.lib.or_set_pressure_solver(self.h, kind)

    def compute_rhs(self, state, config):
        """Explicit RHS (oracle only); caller applies cons->prim + ghost fill first."""
        assert self.impl == "c"
        f, keep = _fieldptrs(state)
        out = {fid: np.zeros(self.n) for fid in state.field_ids() if fid in (0, 1, 2, 3, 4)}
        arr = (C.POINTER(C.c_double) * _abi.N_FIELD_SLOTS)()
        for fid, a in out.items():
            arr[fid] = _abi.dptr(a)
        cc = config.to_c()
        self.lib.or_compute_rhs(self.h, f, C.byref(cc), arr)
        return out


def riemann_flux(impl, rtype, L, R, direction, dim, incP, igr, nph):
    lib, pre = _lib(impl)
    L = np.ascontiguousarray(L, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64)
    out = np.zeros(15)
    getattr(lib, pre + "riemann_flux")(int(rtype), _abi.dptr(L), _abi.dptr(R), int(direction), int(dim), int(incP),
                                       int(igr), int(nph), _abi.dptr(out))
    return out


def recon_c(order, side, v, eps):
    lib, _ = _lib("c")
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib.or_recon(int(order), int(side), _abi.dptr(v), float(eps))


def recon_line_ref(order, eps, nx, ng, v):
    """Left/right rho face states of a 1D line through the reference's public
    Reconstructor API (v: nx + 2 ng padded values)."""
    lib, _ = _lib("ref")
    v = np.ascontiguousarray(v, dtype=np.float64)
    L = np.zeros(nx + 1)
    R = np.zeros(nx + 1)
    lib.ref_recon_line(int(order), float(eps), int(nx), int(ng), _abi.dptr(v), _abi.dptr(L), _abi.dptr(R))
    return L, R
