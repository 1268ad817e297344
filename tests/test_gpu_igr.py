"""GPU parity of IGR (information geometric regularisation): the entropic
pressure sigma from in-place lexicographic Gauss-Seidel sweeps (a tiled
wavefront on the GPU), warm start at step 0, LF flux with sigma, in the
explicit and the semi-implicit step.

Bar: relative L1 and Linf <= 1e-12 after N steps for every field (sigma
included) against the FP64 oracle and against the reference's own output
(igr_ref.npz); identical time steps and Jacobi counts.  All calls go
through the C ABI.
"""
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


@pytest.mark.parametrize("key", list(problems.igr_cases()))
def test_igr_cases(gpu, oracle_c, key):
    d = np.load(os.path.join(GOLDEN, "igr_ref.npz"))
    build, rs, nsteps, semi = problems.igr_cases()[key]
    cfg, mesh, st, eos = build()
    rsolver = RS[rs](eos)
    sg, sc = st.copy(), st.copy()
    if semi:
        gpu_s = SemiImplicitSolver(mesh, rsolver, JacobiPressureSolver(), eos, None, cfg)
    else:
        gpu_s = ExplicitSolver(mesh, rsolver, eos, None, cfg)
    cpu = oracle_c.CpuSolver(mesh, cfg, rs, eos, "c")
    gpu_s.upload(sg)
    for k in range(nsteps):
        cfg.step = k
        dg = gpu_s.step_resident(cfg)
        if semi:
            dc, ic = cpu.semi_step(sc, cfg)
            assert gpu_s.lastPressureIterations() == ic == d[f"{key}_iters"][k]
        else:
            dc = cpu.explicit_step(sc, cfg)
        assert abs(dg - dc) <= 1e-13 * dc and abs(dg - d[f"{key}_dts"][k]) <= 1e-13 * dc
    gpu_s.download(sg)
    gpu_s.close()
    for fid in sc.field_ids():
        ref, orc, got = d[f"{key}_f{fid}"], mesh.interior(sc.field(fid)), mesh.interior(sg.field(fid))
        b1, bi = noise_floor_bar(ref, orc)
        for other in (orc, ref):
            l1, li = rel_errs(got, other)
            assert l1 <= b1 and li <= bi, f"field {fid}: rel L1 {l1:.3e} (bar {b1:.1e}), Linf {li:.3e} (bar {bi:.1e})"


def test_igr_resident_advance_matches_steps(gpu):
    """IGR in the device-resident time loop (warm start at the first step,
    IGRIters afterwards) is bitwise the sequence of single steps."""
    build, rs, nsteps, _ = problems.igr_cases()["e2_lf"]
    cfg, mesh, st, eos = build()
    a = ExplicitSolver(mesh, LFSolver(eos), eos, None, cfg)
    b = ExplicitSolver(mesh, LFSolver(eos), eos, None, cfg)
    a.upload(st)
    b.upload(st)
    t = 0.0
    for k in range(4):
        cfg.step = k
        t += a.step_resident(cfg)
    cfg.step = 0
    tb, nb = b.advance(cfg, 4, 1e9, 0.0)
    assert nb == 4 and abs(t - tb) <= 1e-15 * t
    for fid in (0, 1, 2, 4, 9):
        assert np.array_equal(a.download_field(fid), b.download_field(fid)), fid
    a.close()
    b.close()
