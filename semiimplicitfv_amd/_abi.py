"""ctypes mirror of include/sifv_gpu.h and the loader of the HIP library.

The product path has exactly one implementation: libsifv_hip.so built from
semiimplicitfv_amd/csrc for gfx950.  load_library() raises when it is
missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

MAX_PHASES = 8

# status codes
OK, ERR_INVALID, ERR_HIP, ERR_NONFINITE, ERR_UNSUPPORTED, ERR_NOMEM = range(6)

# field ids (SolutionState members)
F_RHO, F_RHOU, F_RHOV, F_RHOW, F_RHOE = 0, 1, 2, 3, 4
F_VELU, F_VELV, F_VELW, F_PRES, F_SIGMA = 5, 6, 7, 8, 9
F_ALPHARHO = 16
F_ALPHA = 24
N_FIELD_SLOTS = 40

FP64, MIXED = 0, 1
EOS_IDEAL, EOS_STIFFENED = 0, 1
VARSET_CONS, VARSET_PRIM, VARSET_ALL = 0, 1, 2


class PhaseEOS_C(C.Structure):
    _fields_ = [("gamma", C.c_double), ("pInf", C.c_double)]


class _Explicit(C.Structure):
    _fields_ = [("cfl", C.c_double), ("constDt", C.c_double), ("maxDt", C.c_double), ("minDt", C.c_double)]


class _Semi(C.Structure):
    _fields_ = [("cfl", C.c_double), ("maxDt", C.c_double), ("minDt", C.c_double),
                ("maxPressureIters", C.c_int), ("pressureTol", C.c_double)]


class _IGR(C.Structure):
    _fields_ = [("alphaCoeff", C.c_double), ("IGRIters", C.c_int), ("IGRWarmStartIters", C.c_int)]


class _Multi(C.Structure):
    _fields_ = [("nPhases", C.c_int), ("phases", PhaseEOS_C * MAX_PHASES), ("alphaMin", C.c_double)]


class _Body(C.Structure):
    _fields_ = [("a", C.c_double * 3), ("b", C.c_double * 3), ("c", C.c_double * 3), ("d", C.c_double * 3)]


class Config_C(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("nGhost", C.c_int), ("RKOrder", C.c_int), ("useIGR", C.c_int),
        ("semiImplicit", C.c_int), ("reconOrder", C.c_int), ("wenoEps", C.c_double),
        ("step", C.c_int), ("time", C.c_double),
        ("explicitParams", _Explicit), ("semiImplicitParams", _Semi), ("igrParams", _IGR),
        ("multiPhaseParams", _Multi), ("bodyForceParams", _Body),
        ("viscousMu", C.c_double), ("surfaceTensionSigma", C.c_double),
        ("viscousNPhaseMu", C.c_int), ("viscousPhaseMu", C.c_double * MAX_PHASES),
        ("surfaceTensionEps", C.c_double),
    ]


class Mesh_C(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nGhost", C.c_int),
        ("xNodes", C.POINTER(C.c_double)), ("yNodes", C.POINTER(C.c_double)), ("zNodes", C.POINTER(C.c_double)),
        ("bc", C.c_int * 6), ("periods", C.c_int * 3),
    ]


class Eos_C(C.Structure):
    _fields_ = [("kind", C.c_int), ("gamma", C.c_double), ("pInf", C.c_double)]


_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsifv_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "sifv_gpu.h")

_lib = None


class SifvError(RuntimeError):
    """Raised for a non-OK status of the C ABI (the message is sifv_last_error())."""

    def __init__(self, code, msg):
        super().__init__(f"[status {code}] {msg}")
        self.code = code


class InvalidArgument(SifvError, ValueError):
    """Mirror of std::invalid_argument thrown by SimulationConfig::validate."""


def header_symbols(path: str = HEADER_PATH):
    """Names of the functions declared in include/sifv_gpu.h."""
    txt = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sifv_\w+)\s*\(", txt, re.M)))


def _declare(lib):
    P = C.POINTER
    vp = C.c_void_p
    d = C.c_double
    i = C.c_int
    sz = C.c_size_t
    sigs = {
        "sifv_create": (i, [P(Mesh_C), P(Config_C), i, P(Eos_C), i, P(vp)]),
        "sifv_destroy": (i, [vp]),
        "sifv_last_error": (C.c_char_p, []),
        "sifv_abi_version": (i, []),
        "sifv_total_cells": (i, [vp, P(sz)]),
        "sifv_upload": (i, [vp, i, P(d), sz]),
        "sifv_download": (i, [vp, i, P(d), sz]),
        "sifv_explicit_step": (i, [vp, P(Config_C), d, P(d)]),
        "sifv_explicit_advance": (i, [vp, P(Config_C), i, d, P(d), P(i)]),
        "sifv_explicit_advance_noalloc": (i, [vp, P(Config_C), i, d, P(d), P(i)]),
        "sifv_semi_implicit_step": (i, [vp, P(Config_C), d, P(d), P(i)]),
        "sifv_jacobi_solve": (i, [vp, P(d), P(d), P(d), P(d), d, d, i, P(i), P(d)]),
        "sifv_pcg_solve": (i, [vp, P(d), P(d), P(d), P(d), d, d, i, P(i)]),
        "sifv_mgpcg_solve": (i, [vp, P(d), P(d), P(d), P(d), d, d, i, P(i)]),
        "sifv_set_pressure_solver": (i, [vp, i]),
        "sifv_fill_ghosts": (i, [vp, i]),
        "sifv_fill_scalar_ghosts": (i, [vp, i]),
        "sifv_acoustic_dt": (i, [vp, d, d, P(d)]),
        "sifv_advective_dt": (i, [vp, d, d, P(d)]),
        "sifv_cons_to_prim": (i, [vp]),
        "sifv_timing_enable": (i, [vp, i]),
        "sifv_timing_read": (i, [vp, P(d), P(C.c_long), i]),
        "sifv_synchronize": (i, [vp]),
        "sifv_set_tile": (i, [vp, i]),
        "sifv_set_stage_slabs": (i, [vp, i]),
        "sifv_set_graphs": (i, [vp, i]),
        "sifv_graph_replays": (i, [vp, P(C.c_long)]),
        "sifv_write_vtr": (i, [vp, C.c_char_p, i]),
        "sifv_write_pvtr": (i, [vp, C.c_char_p, C.c_char_p]),
        "sifv_write_pvd": (i, [C.c_char_p, C.c_char_p, d, C.c_char_p]),
        "sifv_adjoint_forward": (i, [vp, P(Config_C), i, i]),
        "sifv_adjoint_backward": (i, [vp, vp, vp, P(d)]),
        "sifv_adjoint_backward_host": (i, [vp, P(d), P(d), P(d)]),
        "sifv_state_device_ptr": (i, [vp, P(vp)]),
        "sifv_ensemble_create": (i, [i, P(i), P(d), P(Config_C), i, P(Eos_C), i, P(vp)]),
        "sifv_ensemble_destroy": (i, [vp]),
        "sifv_ensemble_total_cells": (i, [vp, P(sz)]),
        "sifv_ensemble_upload": (i, [vp, i, P(d), sz]),
        "sifv_ensemble_download": (i, [vp, i, P(d), sz]),
        "sifv_ensemble_set_shape": (i, [vp, i, i]),
        "sifv_ensemble_run": (i, [vp, d, i, P(i), P(d), P(d)]),
        "sifv_ensemble_autotune": (i, [vp, d, i, P(i), P(i), P(d)]),
    }
    for name, (res, args) in sigs.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
    return lib


def load_library(path: str | None = None):
    """Load libsifv_hip.so (built by semiimplicitfv_amd.build).  Raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"HIP library {p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(there is no CPU fallback)")
    lib = _declare(C.CDLL(p))
    if path is None:
        _lib = lib
    return lib


def check(rc: int):
    if rc == OK:
        return
    msg = load_library().sifv_last_error().decode(errors="replace")
    if rc == ERR_INVALID:
        raise InvalidArgument(rc, msg)
    raise SifvError(rc, msg)


def dptr(a):
    """double* of a contiguous float64 numpy array."""
    return a.ctypes.data_as(C.POINTER(C.c_double))
