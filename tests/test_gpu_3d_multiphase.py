"""GPU parity of the explicit path in 3D (x+y pass and z pass) and with
two-phase states (upwind alpha-rho fluxes, alpha div(u) source, post-update
clamps, mixture EOS), against the oracle.

Bar: relative L1 and Linf <= 1e-12 after N steps against the FP64 CPU
oracle on the same inputs.  Water/air states are ill-conditioned (the
stiffened-gas energy carries gamma*pInf ~ 2.6e9 beside pressures ~ 1e5),
and there the compiled reference and the oracle already differ by up to
~2e-11 through rounding alone; for those the bar of each field is
max(1e-12, 3 x that measured reference-vs-oracle distance), applied against
both the oracle and the reference's own output.  All calls go through the
C ABI.
"""
import itertools
import os

import numpy as np
import pytest

from semiimplicitfv_amd import (BoundaryCondition, ExplicitSolver, HLLCSolver, IdealGasEOS, LFSolver, RusanovSolver,
                                _abi)
from tests.golden import problems
from tests.helpers import ALL_BCS, make_case, make_two_phase_case, noise_floor_bar, rel_errs

pytestmark = pytest.mark.gpu

TOL = 1e-12
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
RS = {"hllc": HLLCSolver, "rusanov": RusanovSolver, "lf": LFSolver}


def _pair(oracle, cfg, mesh, st, eos, rsolver, nsteps, rows=0, slabs=0):
    sg, sc = st.copy(), st.copy()
    gpu = ExplicitSolver(mesh, rsolver, eos, None, cfg)
    if rows:
        gpu.set_tile(rows)
    if slabs:
        gpu.set_stage_slabs(slabs)
    cpu = oracle.CpuSolver(mesh, cfg, rsolver.kind(), eos, "c")
    gpu.upload(sg)
    dts = []
    for s in range(nsteps):
        cfg.step = s
        dts.append((gpu.step_resident(cfg), cpu.explicit_step(sc, cfg)))
    gpu.download(sg)
    gpu.close()
    return sg, sc, dts


def _check(mesh, sg, sc, dts, tol=TOL):
    for dg, dc in dts:
        assert abs(dg - dc) <= 1e-13 * dc, (dg, dc)
    for fid in sc.field_ids():
        l1, li = rel_errs(mesh.interior(sg.field(fid)), mesh.interior(sc.field(fid)))
        assert l1 <= tol and li <= tol, f"field {fid}: rel L1 {l1:.3e}, rel Linf {li:.3e}"


@pytest.mark.parametrize("recon,rsname", [(2, "hllc"), (2, "rusanov"), (2, "lf"), (1, "hllc"), (0, "rusanov"),
                                          (5, "hllc"), (4, "lf")])
def test_3d_recon_riemann(gpu, oracle_c, recon, rsname):
    cfg, mesh, st, eos = make_case(3, (14, 11, 9), recon=recon, seed=recon + 3, periods=(1, 0, 1))
    sg, sc, dts = _pair(oracle_c, cfg, mesh, st, eos, RS[rsname](eos), 4)
    _check(mesh, sg, sc, dts)


@pytest.mark.parametrize("bx,bz", list(itertools.product(ALL_BCS[:4], ALL_BCS[:4]))[::3])
def test_3d_boundary_rules(gpu, oracle_c, bx, bz):
    bcs = [bx, ALL_BCS[(bx + 1) % 4], BoundaryCondition.Symmetry, BoundaryCondition.NoSlipWall, bz,
           ALL_BCS[(bz + 2) % 4]]
    cfg, mesh, st, eos = make_case(3, (9, 8, 10), bcs=bcs, seed=7 + bx + bz, nonuniform=True)
    sg, sc, dts = _pair(oracle_c, cfg, mesh, st, eos, HLLCSolver(eos), 3)
    _check(mesh, sg, sc, dts)


@pytest.mark.parametrize("rows", [0, 1, 3])
def test_3d_multi_strip_segments(gpu, oracle_c, rows):
    """nx > one 62-column strip, several y and z segments, RK2."""
    cfg, mesh, st, eos = make_case(3, (70, 19, 21), rk=2, seed=11, periods=(0, 1, 0), nonuniform=True)
    sg, sc, dts = _pair(oracle_c, cfg, mesh, st, eos, HLLCSolver(eos), 2, rows=rows)
    _check(mesh, sg, sc, dts)


@pytest.mark.parametrize("slabs", [2, 3, 8])
def test_3d_stage_slabs(gpu, oracle_c, slabs):
    """A stage split into slabs of z segments (x+y pass of the next slab on one
    stream beside the z pass of this one on another): 7 z segments of 3 planes
    cut 3+4, 2+2+3 and into 7 one-segment slabs (8 capped at the segment
    count), RK3 with the fused dt partials.  Same bits as one launch of each
    pass, and the oracle's bar."""
    cfg, mesh, st, eos = make_case(3, (70, 19, 21), rk=3, seed=13, periods=(0, 1, 0), nonuniform=True)
    sg, sc, dts = _pair(oracle_c, cfg, mesh, st, eos, HLLCSolver(eos), 3, rows=3, slabs=slabs)
    _check(mesh, sg, sc, dts)
    s1, _, dts1 = _pair(oracle_c, cfg, mesh, st, eos, HLLCSolver(eos), 3, rows=3, slabs=1)
    assert [d for d, _ in dts] == [d for d, _ in dts1]
    for fid in sc.field_ids():
        assert np.array_equal(mesh.interior(sg.field(fid)), mesh.interior(s1.field(fid))), f"field {fid}"


def test_3d_stage_slabs_device_loop(gpu):
    """The device-resident time loop (sifv_explicit_advance: dt, time and the
    active flag on the device, every launch checking it) with 3D stages in 3
    slabs: the same times, step counts and bits as one launch of each pass,
    across calls that end mid-segment at an end time."""
    cfg, mesh, st, eos = make_case(3, (40, 18, 24), rk=3, seed=17, periods=(1, 0, 1), nonuniform=True)
    solvers = []
    for slabs in (3, 1):
        g = ExplicitSolver(mesh, HLLCSolver(eos), eos, None, cfg)
        g.set_tile(4)   # 6 z segments of 4 planes: 3 slabs of 2
        g.set_stage_slabs(slabs)
        g.upload(st)
        solvers.append(g)
    probe = ExplicitSolver(mesh, HLLCSolver(eos), eos, None, cfg)
    probe.upload(st)
    t5, _ = probe.advance(cfg, 5, 1e30, 0.0)
    probe.close()
    t_end = 3.3 * t5   # after 5 + 9 steps (~2.8 t5) the third call reaches it mid-segment
    hist = [[], []]
    for nsteps in (5, 9, 6):
        for q, g in enumerate(solvers):
            t0 = hist[q][-1][0] if hist[q] else 0.0
            hist[q].append(g.advance(cfg, nsteps, t_end, t0))
    assert abs(hist[0][-1][0] - t_end) <= 1e-12 * t_end and 0 < hist[0][-1][1] < 6, hist
    assert hist[0] == hist[1], hist
    for fid in st.field_ids():
        a, b = solvers[0].download_field(fid), solvers[1].download_field(fid)
        assert np.array_equal(mesh.interior(a), mesh.interior(b)), fid
    for g in solvers:
        g.close()


def test_stage_slabs_range(gpu):
    """sifv_set_stage_slabs takes 0 (the library's choice) to 8 and refuses the rest."""
    cfg, mesh, st, eos = make_case(3, (16, 12, 10), seed=2)
    g = ExplicitSolver(mesh, HLLCSolver(eos), eos, None, cfg)
    for ok in (0, 1, 8):
        g.set_stage_slabs(ok)
    for bad in (-1, 9):
        with pytest.raises(_abi.SifvError):
            g.set_stage_slabs(bad)
    g.close()


def test_3d_periodic_vortex_tube_conservation(gpu):
    """Fully periodic 3D: mass, momentum and energy conserved to round-off."""
    cfg, mesh, st, eos = make_case(3, (32, 32, 16), periods=(1, 1, 1), shock=False)
    g = ExplicitSolver(mesh, HLLCSolver(eos), eos, None, cfg)
    g.upload(st)
    before = [mesh.interior(st.field(f)).sum() for f in (0, 1, 2, 3, 4)]
    for s in range(5):
        g.step_resident(cfg)
    after = [mesh.interior(g.download_field(f)).sum() for f in (0, 1, 2, 3, 4)]
    for b, a in zip(before, after):
        assert abs(a - b) <= 1e-12 * max(1.0, abs(b)), (a, b)
    g.close()


def test_liquid_gas_tube_golden(gpu, oracle_c):
    """Water/air shock tube, 60 steps: oracle and the reference's own output."""
    d = np.load(os.path.join(GOLDEN, "liquid_gas_1d.npz"))
    cfg, mesh, st, eos, rs = problems.liquid_gas_tube(200)
    sg, sc, dts = _pair(oracle_c, cfg, mesh, st, eos, rs, 60)
    for k, (dg, dc) in enumerate(dts):
        assert abs(dg - dc) <= 1e-13 * dc and abs(dg - d["dts"][k]) <= 1e-12 * d["dts"][k]
    for fid in sc.field_ids():
        ref, orc, got = mesh.interior(d[f"f{fid}"]), mesh.interior(sc.field(fid)), mesh.interior(sg.field(fid))
        b1, bi = noise_floor_bar(ref, orc)
        for other in (orc, ref):
            l1, li = rel_errs(got, other)
            assert l1 <= b1 and li <= bi, f"field {fid}: rel L1 {l1:.3e} (bar {b1:.1e}), Linf {li:.3e} (bar {bi:.1e})"


@pytest.mark.parametrize("key", ["x1_rs2", "x1_rs1", "x2_rs2", "x2_rs0", "x3_rs2", "x3_rs1"])
def test_two_phase_explicit(gpu, oracle_c, key):
    """Seeded water/air states, 5 steps; bar per field: max(1e-12, 3x the
    reference-vs-oracle distance), against the oracle and against the
    reference's own output."""
    d = np.load(os.path.join(GOLDEN, "two_phase_ref.npz"))
    build, rs, nsteps, _ = problems.two_phase_cases()[key]
    cfg, mesh, st, eos = build()
    rsolver = {0: LFSolver, 1: RusanovSolver, 2: HLLCSolver}[rs](eos)
    sg, sc, dts = _pair(oracle_c, cfg, mesh, st, eos, rsolver, nsteps)
    for k, (dg, dc) in enumerate(dts):
        assert abs(dg - dc) <= 1e-13 * dc and abs(dg - d[f"{key}_dts"][k]) <= 1e-13 * dc
    for fid in sc.field_ids():
        ref, orc, got = d[f"{key}_f{fid}"], mesh.interior(sc.field(fid)), mesh.interior(sg.field(fid))
        b1, bi = noise_floor_bar(ref, orc)
        for other in (orc, ref):
            l1, li = rel_errs(got, other)
            assert l1 <= b1 and li <= bi, f"field {fid}: rel L1 {l1:.3e} (bar {b1:.1e}), Linf {li:.3e} (bar {bi:.1e})"


def test_two_phase_ghost_fill_and_prim(gpu, oracle_c):
    """Ghost fills of two-phase states (every rule, periods axis) bitwise."""
    for dim in (1, 2, 3):
        cfg, mesh, st, bcs, periods = problems.random_ghost_case(dim, nph=2)
        eos = IdealGasEOS(1.4, 1.0, cfg)
        g = ExplicitSolver(mesh, HLLCSolver(eos), eos, None, cfg)
        for vs in (0, 1, 2):
            g.upload(st)
            g.fill_ghosts(vs)
            ref = st.copy()
            oracle_c.CpuSolver(mesh, cfg, 2, eos, "c").fill_ghosts(ref, vs)
            for fid in ref.field_ids():
                assert np.array_equal(g.download_field(fid), ref.field(fid)), (dim, vs, fid)
        g.close()


def test_3d_dt_and_cons_to_prim(gpu, oracle_c):
    cfg, mesh, st, eos = make_two_phase_case(3, (12, 9, 7), seed=5, nonuniform=True)
    g = ExplicitSolver(mesh, HLLCSolver(eos), eos, None, cfg)
    g.upload(st)
    cpu = oracle_c.CpuSolver(mesh, cfg, 2, eos, "c")
    a, b = g.acoustic_dt(0.6, 10.0), cpu.acoustic_dt(st, cfg, 0.6, 10.0)
    assert abs(a - b) <= 1e-15 * b
    a, b = g.advective_dt(0.8, 10.0), cpu.advective_dt(st, 0.8, 10.0)
    assert abs(a - b) <= 1e-15 * b
    g.cons_to_prim()
    cpu.cons_to_prim(st, cfg)
    for fid in (_abi.F_RHO, _abi.F_VELU, _abi.F_VELW, _abi.F_PRES):
        l1, li = rel_errs(mesh.interior(g.download_field(fid)), mesh.interior(st.field(fid)))
        assert li <= 1e-12 * (4.4 * 6.0e8 / 1e5 if fid == _abi.F_PRES else 1.0), (fid, li)
    g.close()


def test_mixed_precision_3d(gpu, oracle_c):
    """FP32 bulk / FP64 EOS form in 3D (x+y pass and z pass in float):
    relative L1 error vs the FP64 oracle <= 1e-5 after 6 smooth steps."""
    cfg, mesh, st, eos = make_case(3, (40, 24, 20), periods=(1, 1, 1), shock=False, seed=3)
    sg, sc = st.copy(), st.copy()
    gpu = ExplicitSolver(mesh, HLLCSolver(eos), eos, None, cfg, precision="mixed")
    cpu = oracle_c.CpuSolver(mesh, cfg, 2, eos, "c")
    gpu.upload(sg)
    for s in range(6):
        cfg.step = s
        dg, dc = gpu.step_resident(cfg), cpu.explicit_step(sc, cfg)
        assert abs(dg - dc) <= 1e-5 * dc
    gpu.download(sg)
    gpu.close()
    for fid in (0, 1, 2, 3, 4):
        l1, li = rel_errs(mesh.interior(sg.field(fid)), mesh.interior(sc.field(fid)))
        assert l1 <= 1e-5, (fid, l1, li)
