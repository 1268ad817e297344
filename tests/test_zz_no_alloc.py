"""sifv_explicit_advance_noalloc (the allocation-free time loop) against
sifv_explicit_advance, both called through the C ABI on handles built from
the same inputs.  Both run the same kernels; the allocation-free form keeps
its dt partials in the library's static device array and takes the first dt
of each call from the primitive arrays, which yields the same cell minima, so
the results are bitwise equal.  Two calls in a row, the second on a smaller
mesh, then a hand-over between the two entry points on one pair of handles.
"""
import ctypes as C

import numpy as np
import pytest

from semiimplicitfv_amd import ExplicitSolver, HLLCSolver, _abi, cases, uniform_mesh
from tests.helpers import make_case

pytestmark = pytest.mark.gpu

FIELDS2 = (_abi.F_RHO, _abi.F_RHOU, _abi.F_RHOV, _abi.F_RHOE, _abi.F_PRES)


def _noalloc(g, cfg, nsteps, end=1e30, t0=0.0):
    lib = _abi.load_library()
    t = C.c_double(t0)
    s = C.c_int()
    cc = cfg.to_c()
    rc = lib.sifv_explicit_advance_noalloc(g._h, C.byref(cc), int(nsteps), float(end), C.byref(t), C.byref(s))
    return rc, t.value, s.value


def _pair(build, precision="fp64"):
    cfg, mesh, st, eos, rs = build()
    a = ExplicitSolver(mesh, rs, eos, None, cfg, precision=precision)
    b = ExplicitSolver(mesh, rs, eos, None, cfg, precision=precision)
    a.upload(st)
    b.upload(st)
    return cfg, mesh, a, b


def _same(a, b, fids):
    for f in fids:
        x, y = a.download_field(f), b.download_field(f)
        assert np.isfinite(x).all()
        assert np.array_equal(x, y), f


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_advance_two_calls(gpu, precision):
    """First call on a 192^2 vortex, second on a 96^2 vortex (smaller input):
    the allocation-free loop equals sifv_explicit_advance bitwise."""
    for n, steps in ((192, 12), (96, 7)):
        cfg, mesh, a, b = _pair(lambda: cases.isentropic_vortex_2d(n), precision)
        ta, na = a.advance(cfg, steps, 1e30, 0.0)
        rc, tb, nb = _noalloc(b, cfg, steps)
        assert rc == 0, _abi.load_library().sifv_last_error()
        assert (ta, na) == (tb, nb) and na == steps
        _same(a, b, FIELDS2)
        a.close()
        b.close()


def test_noalloc_hand_over_and_end_time(gpu):
    """Regular -> allocation-free -> regular on one handle equals regular
    throughout on the other, including a call that stops at an end time."""
    cfg, mesh, a, b = _pair(lambda: cases.isentropic_vortex_2d(128))
    t_a, n_a = a.advance(cfg, 5, 1e30, 0.0)
    t_b, n_b = b.advance(cfg, 5, 1e30, 0.0)
    t_a, n_a = a.advance(cfg, 6, 1e30, t_a)
    rc, t_b, n_b = _noalloc(b, cfg, 6, 1e30, t_b)
    assert rc == 0 and (t_a, n_a) == (t_b, n_b)
    _same(a, b, FIELDS2)
    end = t_a + 2.5 * 0.03
    t_a, n_a = a.advance(cfg, 50, end, t_a)
    t_b, n_b = b.advance(cfg, 50, end, t_b)
    assert (t_a, n_a) == (t_b, n_b) and t_a == end
    _same(a, b, FIELDS2)
    rc, t_b, n_b = _noalloc(b, cfg, 4, 1e30, t_b)
    t_a, n_a = a.advance(cfg, 4, 1e30, t_a)
    assert rc == 0 and (t_a, n_a) == (t_b, n_b)
    _same(a, b, FIELDS2)
    a.close()
    b.close()


def test_noalloc_3d(gpu):
    """3D (x+y pass and z pass, partials from the z pass), smaller second call."""
    for n, steps in (((30, 22, 18), 4), ((20, 14, 12), 3)):
        def build():
            cfg, mesh, st, eos = make_case(3, n, periods=(1, 0, 1), seed=5)
            return cfg, mesh, st, eos, HLLCSolver(eos)
        cfg, mesh, a, b = _pair(build)
        ta, na = a.advance(cfg, steps, 1e30, 0.0)
        rc, tb, nb = _noalloc(b, cfg, steps)
        assert rc == 0 and (ta, na) == (tb, nb)
        _same(a, b, (_abi.F_RHO, _abi.F_RHOU, _abi.F_RHOW, _abi.F_RHOE))
        a.close()
        b.close()


def test_noalloc_refusals(gpu):
    """A mesh that needs more dt partials than the static array holds is
    refused before anything runs (the handle stays usable); so is a call with
    kernel timing on (timing creates events)."""
    lib = _abi.load_library()
    cfg, mesh, st, eos, rs = cases.sod_1d(62 * 66000)
    g = ExplicitSolver(mesh, rs, eos, None, cfg)
    g.upload(st)
    rc, _, _ = _noalloc(g, cfg, 1)
    assert rc == _abi.ERR_INVALID and "65536" in lib.sifv_last_error().decode()
    t, n = g.advance(cfg, 1, 1e30, 0.0)
    assert n == 1 and np.isfinite(g.download_field(_abi.F_RHO)).all()
    g.close()
    cfg, mesh, a, b = _pair(lambda: cases.isentropic_vortex_2d(32))
    b.timing(True)
    rc, _, _ = _noalloc(b, cfg, 1)
    assert rc == _abi.ERR_UNSUPPORTED
    b.timing(False)
    a.close()
    b.close()
