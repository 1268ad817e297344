"""Viscous stress on the GPU path (visc_kernel + the stage update, and the
viscous limit in the time-step reduction) against the oracle and the
reference's own output (tests/golden/viscous_ref.npz): explicit scalar mu
in 1D-3D with no-slip, slip, outflow and periodic faces on stretched meshes,
per-phase mu on a water/air state, and the semi-implicit step (2D/3D scalar
mu with the viscous limit binding, water/air with per-phase mu under
gravity), with the config time advanced per step as runTimeLoop does.  A
case with a large viscosity makes the viscous limit the binding one on the
explicit path; the device-resident time loop equals the host loop of steps
bit for bit; the FP32-bulk form stays within 1e-5 of FP64.

Bars: scalar-mu cases relative L1 and Linf <= 1e-12 against the oracle and
the fixture; the water/air cases use the measured reference-vs-oracle floor
(tests/helpers.py::noise_floor_bar); Jacobi iteration counts equal."""
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


def _build(key, mu=None):
    build, rs, nsteps, semi = problems.viscous_cases()[key]
    cfg, mesh, st, eos = build()
    if mu is not None:
        cfg.viscousParams.mu = mu
    return cfg, mesh, st, eos, rs, nsteps, semi


def _gpu_run(key, mu=None, precision="fp64", its=None):
    cfg, mesh, st, eos, rs, nsteps, semi = _build(key, mu)
    g = (SemiImplicitSolver(mesh, RS[rs](eos), JacobiPressureSolver(), eos, None, cfg) if semi
         else ExplicitSolver(mesh, RS[rs](eos), eos, None, cfg, precision=precision))
    g.upload(st)
    dts, t = [], 0.0
    for k in range(nsteps):
        cfg.step, cfg.time = k, t
        dts.append(g.step_resident(cfg))
        if its is not None:
            its.append(g.lastPressureIterations())
        t += dts[-1]
    g.download(st)
    g.close()
    return mesh, st, dts


def _oracle_run(po, key, mu=None):
    cfg, mesh, st, eos, rs, nsteps, semi = _build(key, mu)
    s = po.CpuSolver(mesh, cfg, rs, eos, "c")
    dts, t = [], 0.0
    for k in range(nsteps):
        cfg.step, cfg.time = k, t
        dts.append(s.semi_step(st, cfg)[0] if semi else s.explicit_step(st, cfg))
        t += dts[-1]
    return st, dts


@pytest.mark.parametrize("key", list(problems.viscous_cases()))
def test_viscous_cases(gpu, oracle_c, key):
    d = np.load(os.path.join(GOLDEN, "viscous_ref.npz"))
    its = [] if problems.viscous_cases()[key][3] else None
    mesh, sg, dts = _gpu_run(key, its=its)
    sc, dtc = _oracle_run(oracle_c, key)
    assert np.allclose(dts, d[f"{key}_dts"], rtol=1e-13, atol=0.0)
    if its is not None:
        assert its == list(d[f"{key}_iters"])
    # water/air: the states carry the case's rounding floor into the dt
    assert np.allclose(dts, dtc, rtol=1e-13 if key.endswith("p") else 1e-14, atol=0.0)
    for fid in sg.field_ids():
        g, c, r = mesh.interior(sg.field(fid)), mesh.interior(sc.field(fid)), d[f"{key}_f{fid}"]
        b1, bi = noise_floor_bar(r, c) if key.endswith("p") else (1e-12, 1e-12)
        for other in (c, r):
            l1, li = rel_errs(g, other)
            assert l1 <= b1 and li <= bi, (key, fid, l1, li, b1, bi)


@pytest.mark.parametrize("key", ["v2", "v3"])
def test_viscous_limit_binds(gpu, oracle_c, key):
    """mu large enough that dx^2 / (2 dim nu) is below the acoustic limit."""
    mu = 2.0
    mesh, sg, dts = _gpu_run(key, mu)
    sc, dtc = _oracle_run(oracle_c, key, mu)
    _, _, dts0 = _gpu_run(key, 0.0)
    assert dts[0] < 0.5 * dts0[0], (dts[0], dts0[0])
    assert np.allclose(dts, dtc, rtol=1e-14, atol=0.0)
    for fid in sg.field_ids():
        l1, li = rel_errs(mesh.interior(sg.field(fid)), mesh.interior(sc.field(fid)))
        assert l1 <= 1e-12 and li <= 1e-12, (key, fid, l1, li)


def test_viscous_device_loop_matches_steps(gpu):
    """advance() runs the viscous kernel and limit inside the device loop."""
    cfg, mesh, st, eos, rs, _, _ = _build("v2")
    a = ExplicitSolver(mesh, RS[rs](eos), eos, None, cfg)
    b = ExplicitSolver(mesh, RS[rs](eos), eos, None, cfg)
    a.upload(st)
    b.upload(st)
    t_end = 0.05
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


def test_viscous_mixed_precision(gpu):
    """FP32-bulk form with viscosity: within 1e-5 (relative L1) of FP64."""
    m64, s64, _ = _gpu_run("v2", precision="fp64")
    m32, s32, _ = _gpu_run("v2", precision="mixed")
    for fid in (1, 2, 4):
        l1, _ = rel_errs(m32.interior(s32.field(fid)), m64.interior(s64.field(fid)))
        assert l1 <= 1e-5, (fid, l1)
