"""Surface tension on the GPU path (capillary stress in visc_kernel, the
capillary limit in the time-step reduction) against the oracle and the
reference's own output (tests/golden/capillary_ref.npz): water/air states
with planar and circular interfaces, a regularised |grad alpha|, per-phase
viscosity and gravity alongside, on the explicit (1D-3D) and semi-implicit
(2D/3D) solvers, with the config time advanced per step as runTimeLoop does.
The capillary limit binds in c1 (explicit) and cs3 (semi-implicit).  The
device-resident time loop equals the host loop of steps bit for bit.

Bars: the measured reference-vs-oracle floor of each field
(tests/helpers.py::noise_floor_bar), against both the oracle and the fixture;
time steps to 1e-11 relative; Jacobi iteration counts equal."""
import os

import numpy as np
import pytest

from semiimplicitfv_amd import (ExplicitSolver, HLLCSolver, JacobiPressureSolver, LFSolver, RusanovSolver,
                                SemiImplicitSolver)
from tests.golden import problems
from tests.helpers import noise_floor_bar, rel_errs

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
RS = {0: LFSolver, 1: RusanovSolver, 2: HLLCSolver}


def _gpu_run(key, its):
    build, rs, nsteps, semi = problems.capillary_cases()[key]
    cfg, mesh, st, eos = build()
    g = (SemiImplicitSolver(mesh, RS[rs](eos), JacobiPressureSolver(), eos, None, cfg) if semi
         else ExplicitSolver(mesh, RS[rs](eos), eos, None, cfg))
    g.upload(st)
    dts, t = [], 0.0
    for k in range(nsteps):
        cfg.step, cfg.time = k, t
        dts.append(g.step_resident(cfg))
        if semi:
            its.append(g.lastPressureIterations())
        t += dts[-1]
    g.download(st)
    g.close()
    return mesh, st, dts


def _oracle_run(po, key):
    build, rs, nsteps, semi = problems.capillary_cases()[key]
    cfg, mesh, st, eos = build()
    s = po.CpuSolver(mesh, cfg, rs, eos, "c")
    t = 0.0
    for k in range(nsteps):
        cfg.step, cfg.time = k, t
        t += s.semi_step(st, cfg)[0] if semi else s.explicit_step(st, cfg)
    return st


@pytest.mark.parametrize("key", list(problems.capillary_cases()))
def test_capillary_cases(gpu, oracle_c, key):
    d = np.load(os.path.join(GOLDEN, "capillary_ref.npz"))
    its = []
    mesh, sg, dts = _gpu_run(key, its)
    sc = _oracle_run(oracle_c, key)
    assert np.allclose(dts, d[f"{key}_dts"], rtol=1e-11, atol=0.0)
    assert its == list(d[f"{key}_iters"])
    for fid in sg.field_ids():
        g, c, r = mesh.interior(sg.field(fid)), mesh.interior(sc.field(fid)), d[f"{key}_f{fid}"]
        b1, bi = noise_floor_bar(r, c)
        for other in (c, r):
            l1, li = rel_errs(g, other)
            assert l1 <= b1 and li <= bi, (key, fid, l1, li, b1, bi)


def test_capillary_device_loop_matches_steps(gpu):
    """advance() runs the capillary stress and limit inside the device loop."""
    build, rs, _, _ = problems.capillary_cases()["c2"]
    cfg, mesh, st, eos = build()
    a = ExplicitSolver(mesh, RS[rs](eos), eos, None, cfg)
    b = ExplicitSolver(mesh, RS[rs](eos), eos, None, cfg)
    a.upload(st)
    b.upload(st)
    t, n = a.advance(cfg, 9, 1e30, 0.0)
    tb = 0.0
    for _ in range(n):
        cfg.time = tb
        tb += b.step_resident(cfg)
    assert n == 9 and abs(t - tb) <= 1e-15 * max(1.0, t)
    for fid in st.field_ids():
        assert np.array_equal(a.download_field(fid), b.download_field(fid)), fid
    a.close()
    b.close()
