"""Body force a + b cos(c t + d) on the GPU path (explicit 1D-3D and the
semi-implicit two-phase step under gravity) against the oracle and the
reference's own output (tests/golden/body_force_ref.npz), with the config
time advanced per step as runTimeLoop does; the device-resident time loop
(sifv_explicit_advance, acceleration evaluated on the device from the
device time) equals the host loop of steps bit for bit.

Bars: explicit cases relative L1 and Linf <= 1e-12 against the oracle and
the fixture; the water/air semi-implicit case uses the measured
reference-vs-oracle floor (tests/helpers.py::noise_floor_bar)."""
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


def _gpu_run(key):
    build, rs, nsteps, semi = problems.body_force_cases()[key]
    cfg, mesh, st, eos = build()
    g = (SemiImplicitSolver(mesh, RS[rs](eos), JacobiPressureSolver(), eos, None, cfg) if semi
         else ExplicitSolver(mesh, RS[rs](eos), eos, None, cfg))
    g.upload(st)
    dts, its, t = [], [], 0.0
    for k in range(nsteps):
        cfg.step, cfg.time = k, t
        dt = g.step_resident(cfg)
        dts.append(dt)
        if semi:
            its.append(g.lastPressureIterations())
        t += dt
    g.download(st)
    g.close()
    return mesh, st, dts, its


def _oracle_run(po, key):
    build, rs, nsteps, semi = problems.body_force_cases()[key]
    cfg, mesh, st, eos = build()
    s = po.CpuSolver(mesh, cfg, rs, eos, "c")
    t = 0.0
    for k in range(nsteps):
        cfg.step, cfg.time = k, t
        t += s.semi_step(st, cfg)[0] if semi else s.explicit_step(st, cfg)
    return st


@pytest.mark.parametrize("key", list(problems.body_force_cases()))
def test_body_force_cases(gpu, oracle_c, key):
    d = np.load(os.path.join(GOLDEN, "body_force_ref.npz"))
    mesh, sg, dts, its = _gpu_run(key)
    sc = _oracle_run(oracle_c, key)
    assert np.allclose(dts, d[f"{key}_dts"], rtol=1e-13, atol=0.0)
    assert list(its) == list(d[f"{key}_iters"])
    for fid in sg.field_ids():
        g, c, r = mesh.interior(sg.field(fid)), mesh.interior(sc.field(fid)), d[f"{key}_f{fid}"]
        b1, bi = noise_floor_bar(r, c) if key.startswith("s") else (1e-12, 1e-12)
        for other in (c, r):
            l1, li = rel_errs(g, other)
            assert l1 <= b1 and li <= bi, (key, fid, l1, li, b1, bi)


def test_body_force_device_loop_matches_steps(gpu):
    """advance() evaluates the acceleration from the device-resident time."""
    build, rs, nsteps, _ = problems.body_force_cases()["e2"]
    cfg, mesh, st, eos = build()
    a = ExplicitSolver(mesh, RS[rs](eos), eos, None, cfg)
    b = ExplicitSolver(mesh, RS[rs](eos), eos, None, cfg)
    a.upload(st)
    b.upload(st)
    t_end = 0.03
    t, n = a.advance(cfg, 100000, t_end, 0.0)
    tb, nb = 0.0, 0
    while tb < t_end:
        cfg.time = tb
        tb += b.step_resident(cfg, t_end - tb)
        nb += 1
    assert n == nb and abs(t - tb) <= 1e-15
    for fid in st.field_ids():
        assert np.array_equal(a.download_field(fid), b.download_field(fid)), fid
    a.close()
    b.close()


def test_body_force_mixed_precision(gpu):
    """FP32-bulk form with a body force: within 1e-5 (relative L1) of FP64."""
    build, rs, nsteps, _ = problems.body_force_cases()["e2"]
    out = []
    for prec in ("fp64", "mixed"):
        cfg, mesh, st, eos = build()
        g = ExplicitSolver(mesh, RS[rs](eos), eos, None, cfg, precision=prec)
        g.upload(st)
        t = 0.0
        for k in range(nsteps):
            cfg.step, cfg.time = k, t
            t += g.step_resident(cfg)
        out.append(mesh.interior(g.download_field(1)).copy())
        g.close()
    l1, _ = rel_errs(out[1], out[0])
    assert l1 <= 1e-5, l1
