"""GPU parity of the semi-implicit step (advective stage kernels, pressure
set-up, Jacobi sweeps, correction) and of the standalone Jacobi solve,
against the oracle and the reference's own golden output.

Bar: relative L1 and Linf <= 1e-12 after N steps against the FP64 CPU
oracle, identical time steps and Jacobi iteration counts.  Water/air
states are ill-conditioned (gamma*pInf ~ 2.6e9 beside pressures ~ 1e5):
the compiled reference and the oracle already differ there by up to ~2e-11
through rounding alone, so each field's bar is max(1e-12, 3 x that
measured reference-vs-oracle distance), applied against both the oracle
and the reference's own output.  All calls go through the C ABI.
"""
import os

import numpy as np
import pytest

from semiimplicitfv_amd import (BoundaryCondition, HLLCSolver, IdealGasEOS, JacobiPressureSolver, LFSolver,
                                RusanovSolver, SemiImplicitSolver, _abi)
from tests.golden import problems
from tests.helpers import make_semi_case, noise_floor_bar, rel_errs

pytestmark = pytest.mark.gpu

TOL = 1e-12
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
RS = {"hllc": HLLCSolver, "rusanov": RusanovSolver, "lf": LFSolver}


def _pair(oracle, cfg, mesh, st, eos, rsolver, nsteps, rows=0, slabs=0):
    sg, sc = st.copy(), st.copy()
    gpu = SemiImplicitSolver(mesh, rsolver, JacobiPressureSolver(), eos, None, cfg)
    if rows:
        gpu.set_tile(rows)
    if slabs:
        gpu.set_stage_slabs(slabs)
    cpu = oracle.CpuSolver(mesh, cfg, rsolver.kind(), eos, "c")
    gpu.upload(sg)
    hist = []
    for s in range(nsteps):
        cfg.step = s
        dg = gpu.step_resident(cfg)
        dc, ic = cpu.semi_step(sc, cfg)
        hist.append((dg, dc, gpu.lastPressureIterations(), ic))
    gpu.download(sg)
    gpu.close()
    return sg, sc, hist


def _check(mesh, sg, sc, hist, tol=TOL):
    for dg, dc, ig, ic in hist:
        assert abs(dg - dc) <= 1e-13 * dc, (dg, dc)
        assert ig == ic, (ig, ic)
    for fid in sc.field_ids():
        l1, li = rel_errs(mesh.interior(sg.field(fid)), mesh.interior(sc.field(fid)))
        assert l1 <= tol and li <= tol, f"field {fid}: rel L1 {l1:.3e}, rel Linf {li:.3e}"


@pytest.mark.parametrize("rk", [1, 2, 3])
def test_semi_1d_rk_orders(gpu, oracle_c, rk):
    cfg, mesh, st, eos = make_semi_case(1, (96,), rk=rk, seed=rk, nonuniform=True,
                                        bcs=[BoundaryCondition.SlipWall, BoundaryCondition.Outflow] + [0] * 4)
    sg, sc, hist = _pair(oracle_c, cfg, mesh, st, eos, HLLCSolver(eos), 6)
    _check(mesh, sg, sc, hist)


@pytest.mark.parametrize("recon,rsname", [(2, "hllc"), (2, "rusanov"), (2, "lf"), (1, "hllc"), (0, "rusanov"),
                                          (5, "lf")])
def test_semi_2d_recon_riemann(gpu, oracle_c, recon, rsname):
    cfg, mesh, st, eos = make_semi_case(2, (40, 33), recon=recon, seed=recon + 5, periods=(1, 0, 0),
                                        bcs=[0, 0, BoundaryCondition.NoSlipWall, BoundaryCondition.Symmetry, 0, 0])
    sg, sc, hist = _pair(oracle_c, cfg, mesh, st, eos, RS[rsname](eos), 4)
    _check(mesh, sg, sc, hist)


def test_semi_2d_multi_strip(gpu, oracle_c):
    cfg, mesh, st, eos = make_semi_case(2, (130, 41), seed=4, periods=(0, 1, 0), nonuniform=True,
                                        bcs=[BoundaryCondition.SlipWall, BoundaryCondition.Outflow, 0, 0, 0, 0])
    sg, sc, hist = _pair(oracle_c, cfg, mesh, st, eos, HLLCSolver(eos), 3, rows=3)
    _check(mesh, sg, sc, hist)


def test_semi_3d_single_phase(gpu, oracle_c):
    cfg, mesh, st, eos = make_semi_case(3, (13, 10, 11), seed=2, periods=(0, 0, 1),
                                        bcs=[BoundaryCondition.SlipWall, BoundaryCondition.Outflow,
                                             BoundaryCondition.Symmetry, BoundaryCondition.NoSlipWall, 0, 0])
    sg, sc, hist = _pair(oracle_c, cfg, mesh, st, eos, HLLCSolver(eos), 3)
    _check(mesh, sg, sc, hist)


@pytest.mark.parametrize("slabs", [2, 4])
def test_semi_3d_stage_slabs(gpu, oracle_c, slabs):
    """Semi-implicit 3D stages split into slabs of z segments (11 planes in
    segments of 2: 6 segments cut 3+3 and 1+2+1+2): the oracle's bar, and the
    same bits and Jacobi counts as one launch of each pass."""
    cfg, mesh, st, eos = make_semi_case(3, (13, 10, 11), seed=2, periods=(0, 0, 1),
                                        bcs=[BoundaryCondition.SlipWall, BoundaryCondition.Outflow,
                                             BoundaryCondition.Symmetry, BoundaryCondition.NoSlipWall, 0, 0])
    sg, sc, hist = _pair(oracle_c, cfg, mesh, st, eos, HLLCSolver(eos), 3, rows=2, slabs=slabs)
    _check(mesh, sg, sc, hist)
    s1, _, hist1 = _pair(oracle_c, cfg, mesh, st, eos, HLLCSolver(eos), 3, rows=2, slabs=1)
    assert [(h[0], h[2]) for h in hist] == [(h[0], h[2]) for h in hist1]
    for fid in sc.field_ids():
        assert np.array_equal(mesh.interior(sg.field(fid)), mesh.interior(s1.field(fid))), f"field {fid}"


def _check_floor(mesh, sg, sc, ref_fields):
    for fid in sc.field_ids():
        ref, orc, got = ref_fields(fid), mesh.interior(sc.field(fid)), mesh.interior(sg.field(fid))
        b1, bi = noise_floor_bar(ref, orc)
        for other in (orc, ref):
            l1, li = rel_errs(got, other)
            assert l1 <= b1 and li <= bi, f"field {fid}: rel L1 {l1:.3e} (bar {b1:.1e}), Linf {li:.3e} (bar {bi:.1e})"


@pytest.mark.parametrize("key", ["s1", "s2", "s3"])
def test_semi_two_phase(gpu, oracle_c, key):
    """Seeded water/air states, 3 semi-implicit steps (1D runs into the
    400-sweep cap): same dt and Jacobi counts as oracle and reference,
    fields within the rounding-floor bar."""
    d = np.load(os.path.join(GOLDEN, "two_phase_ref.npz"))
    build, rs, nsteps, _ = problems.two_phase_cases()[key]
    cfg, mesh, st, eos = build()
    sg, sc, hist = _pair(oracle_c, cfg, mesh, st, eos, HLLCSolver(eos), nsteps)
    for k, (dg, dc, ig, ic) in enumerate(hist):
        assert abs(dg - dc) <= 1e-13 * dc and abs(dg - d[f"{key}_dts"][k]) <= 1e-13 * dc
        assert ig == ic == d[f"{key}_iters"][k]
    _check_floor(mesh, sg, sc, lambda fid: d[f"{key}_f{fid}"])


def test_bubble_3d_golden(gpu, oracle_c):
    """Air bubble in water (unit cube, 12^3, slip walls), 3 semi-implicit
    steps: oracle, and the reference's own output (iteration counts, dt)."""
    d = np.load(os.path.join(GOLDEN, "bubble3d_semi.npz"))
    cfg, mesh, st, eos, rs = problems.bubble_3d(12, perturb=True)
    sg, sc, hist = _pair(oracle_c, cfg, mesh, st, eos, rs, 3)
    for k, (dg, dc, ig, ic) in enumerate(hist):
        assert ig == ic == d["iters"][k]
        assert abs(dg - dc) <= 1e-15 * dc and abs(dg - d["dts"][k]) <= 1e-15 * d["dts"][k]
    _check_floor(mesh, sg, sc, lambda fid: mesh.interior(d[f"f{fid}"]))


def test_bubble_3d_golden_stage_slabs(gpu, oracle_c):
    """The same water/air bubble with 3D stages in 3 slabs of z segments (12
    planes in segments of 2): the reference's iteration counts and dt, the
    oracle and reference fields at the rounding floor, and the same bits as
    one launch of each pass."""
    d = np.load(os.path.join(GOLDEN, "bubble3d_semi.npz"))
    cfg, mesh, st, eos, rs = problems.bubble_3d(12, perturb=True)
    sg, sc, hist = _pair(oracle_c, cfg, mesh, st, eos, rs, 3, rows=2, slabs=3)
    for k, (dg, dc, ig, ic) in enumerate(hist):
        assert ig == ic == d["iters"][k]
        assert abs(dg - dc) <= 1e-15 * dc and abs(dg - d["dts"][k]) <= 1e-15 * d["dts"][k]
    _check_floor(mesh, sg, sc, lambda fid: mesh.interior(d[f"f{fid}"]))
    s1, _, hist1 = _pair(oracle_c, cfg, mesh, st, eos, rs, 3, rows=2, slabs=1)
    assert [(h[0], h[2]) for h in hist] == [(h[0], h[2]) for h in hist1]
    for fid in sc.field_ids():
        assert np.array_equal(mesh.interior(sg.field(fid)), mesh.interior(s1.field(fid))), f"field {fid}"


@pytest.mark.parametrize("tol,maxit", [(1e-6, 1000), (1e-300, 7), (1e-2, 1000)])
def test_jacobi_standalone(gpu, oracle_c, tol, maxit):
    """sifv_jacobi_solve: iteration count, residual history and the whole
    array (ghosts as the reference leaves them) against the oracle."""
    cfg, mesh, rho, rhoc2, rhs, p0, dt = problems.jacobi_problem()
    eos = IdealGasEOS(1.4, 1.0, cfg)
    cfg.semiImplicitParams.maxPressureIters = maxit
    g = SemiImplicitSolver(mesh, HLLCSolver(eos), JacobiPressureSolver(), eos, None, cfg)
    js = JacobiPressureSolver(g)
    pg = p0.copy()
    hg = np.zeros(maxit)
    ig = js.solve(mesh, rho, rhoc2, rhs, pg, dt, tol, maxit, history=hg)
    cpu = oracle_c.CpuSolver(mesh, cfg, 2, eos, "c")
    pc = p0.copy()
    hc = np.zeros(maxit)
    ic = cpu.jacobi(rho, rhoc2, rhs, pc, dt, tol, maxit, hc)
    assert ig == ic
    assert np.abs(hg[:ig] - hc[:ic]).max() <= 1e-9 * max(hc[0], 1e-300) + 1e-300
    l1, li = rel_errs(pg, pc)
    assert l1 <= 1e-14 and li <= 1e-14, (l1, li)
    if tol == 1e-6:
        dgold = np.load(os.path.join(GOLDEN, "jacobi_3d.npz"))
        assert ig == int(dgold["it_tol1e-06"][0])
        l1, li = rel_errs(mesh.interior(pg), mesh.interior(dgold["p_tol1e-06"]))
        assert l1 <= 1e-14 and li <= 1e-14
    g.close()


def _tile_problem(shape, periods, seed):
    """Variable-coefficient pressure system on a stretched mesh whose extents
    cross the sweep tiles' edges: x past one and two 64-column tiles, y not a
    multiple of the tile rows, z past one 16-plane chunk."""
    from semiimplicitfv_amd import RectilinearMesh, SimulationConfig
    rng = np.random.default_rng(seed)
    dim = len(shape)
    cfg = SimulationConfig(dim=dim, nGhost=3, RKOrder=3, semiImplicit=True)
    nodes = []
    for n in shape:
        h = 1.0 + 0.4 * rng.random(n)
        nodes.append(np.concatenate([[0.0], np.cumsum(h / h.sum())]))
    while len(nodes) < 3:
        nodes.append(np.array([0.0, 1.0]))
    mesh = RectilinearMesh(cfg, nodes[0], nodes[1], nodes[2], periods=periods)
    for d in range(dim):
        if not periods[d]:
            mesh.setBoundaryCondition(2 * d, BoundaryCondition.SlipWall)
            mesh.setBoundaryCondition(2 * d + 1, BoundaryCondition.SlipWall)
    N = mesh.totalCells()
    sh = tuple(reversed(shape))
    rho = np.ones(N)
    mesh.interior(rho)[...] = np.where(rng.random(sh) < 0.3, 1000.0, 1.0) * (1.0 + 0.1 * rng.random(sh))
    rho = problems._fill_scalar_np(mesh, rho)
    rhoc2 = np.zeros(N)
    mesh.interior(rhoc2)[...] = 1.4e5 + 2.25e9 * rng.random(sh)
    rhs = np.zeros(N)
    mesh.interior(rhs)[...] = 1e5 + 1e3 * rng.standard_normal(sh)
    p0 = np.zeros(N)
    mesh.interior(p0)[...] = 1e5 + 10.0 * rng.standard_normal(sh)
    return cfg, mesh, rho, rhoc2, rhs, p0, 2e-5


@pytest.mark.parametrize("shape,periods", [((130, 9, 37), (1, 0, 0)), ((70, 6, 17), (0, 1, 1)),
                                           ((129, 67), (0, 1, 0)), ((65, 5), (1, 1, 0))])
def test_jacobi_tiles(gpu, oracle_c, shape, periods):
    """Tiled Jacobi sweeps across tile, halo and plane-chunk edges: iteration
    count, residual history and every cell against the oracle's sweep (the
    same arithmetic, so to rounding), for a 9-sweep solve (bar 1e-14) and a
    converged one of several hundred sweeps, where FMA-contraction differences
    accumulate (bar 1e-12)."""
    cfg, mesh, rho, rhoc2, rhs, p0, dt = _tile_problem(shape, periods, seed=sum(shape))
    eos = IdealGasEOS(1.4, 1.0, cfg)
    g = SemiImplicitSolver(mesh, HLLCSolver(eos), JacobiPressureSolver(), eos, None, cfg)
    js = JacobiPressureSolver(g)
    cpu = oracle_c.CpuSolver(mesh, cfg, 2, eos, "c")
    for tol, maxit in ((1e-300, 9), (5.0, 2000)):
        pg, pc = p0.copy(), p0.copy()
        hg, hc = np.zeros(maxit), np.zeros(maxit)
        ig = js.solve(mesh, rho, rhoc2, rhs, pg, dt, tol, maxit, history=hg)
        ic = cpu.jacobi(rho, rhoc2, rhs, pc, dt, tol, maxit, hc)
        assert ig == ic, (tol, ig, ic)
        assert np.abs(hg[:ig] - hc[:ic]).max() <= 1e-9 * hc[0]
        l1, li = rel_errs(mesh.interior(pg), mesh.interior(pc))
        bar = 1e-14 if maxit < 10 else TOL
        assert l1 <= bar and li <= bar, (tol, l1, li)
    g.close()


@pytest.mark.parametrize("case", ["s2", "s3", "single2d"])
def test_mixed_precision_jacobi(gpu, case):
    """Semi-implicit step with FP32 Jacobi sweeps and FP64 residual
    refinement (precision="mixed") against the FP64 solve: both stop when the
    FP64 Jacobi update falls below pressureTol, so they differ by the solver
    tolerance, not by FP32 rounding: relative L1 of every field <= 1e-8."""
    if case == "single2d":
        cfg, mesh, st, eos = make_semi_case(2, (40, 33), seed=4, periods=(1, 0, 0), tol=1e-8)
    else:
        build, rs, nsteps, _ = problems.two_phase_cases()[case]
        cfg, mesh, st, eos = build()
    out = []
    for prec in ("fp64", "mixed"):
        g = SemiImplicitSolver(mesh, HLLCSolver(eos), JacobiPressureSolver(), eos, None, cfg, precision=prec)
        s = st.copy()
        g.upload(s)
        its = []
        for k in range(3):
            cfg.step = k
            g.step_resident(cfg)
            its.append(g.lastPressureIterations())
        g.download(s)
        g.close()
        out.append((s, its))
    (a, ia), (b, ib) = out
    # every solve outlasts the FP64 opening chunk of 4 sweeps, so the FP32
    # sweeps and FP64 passes run
    assert all(x > 4 for x in ia), ia
    assert all(0 < x <= cfg.semiImplicitParams.maxPressureIters for x in ib)
    for fid in a.field_ids():
        l1, li = rel_errs(mesh.interior(b.field(fid)), mesh.interior(a.field(fid)))
        assert l1 <= 1e-8, (fid, l1, li, ia, ib)


def test_mixed_precision_short_solves(gpu):
    """A mixed-precision solve opens with the FP64 solver's first chunk of 4
    sweeps; the 12^3 water/air bubble at rest converges within it (1-2 sweeps
    per stage), so its steps are the FP64 steps: the same Jacobi counts and
    the same bits."""
    cfg, mesh, st, eos, rs = problems.bubble_3d(12)
    out = []
    for prec in ("fp64", "mixed"):
        g = SemiImplicitSolver(mesh, HLLCSolver(eos), JacobiPressureSolver(), eos, None, cfg, precision=prec)
        s = st.copy()
        g.upload(s)
        its = []
        for k in range(3):
            cfg.step = k
            its.append((g.step_resident(cfg), g.lastPressureIterations()))
        g.download(s)
        g.close()
        out.append((s, its))
    (a, ia), (b, ib) = out
    assert ia == ib and all(x <= 4 for _, x in ia), (ia, ib)
    for fid in a.field_ids():
        assert np.array_equal(mesh.interior(a.field(fid)), mesh.interior(b.field(fid))), fid


@pytest.mark.parametrize("dim,n", [(2, (40, 36)), (3, (16, 14, 12))])
def test_semi_correction_without_sigma_upload(gpu, dim, n):
    """Without IGR, a handle whose sigma was never uploaded leaves sigma out of
    the pressure correction (it is zero); its steps equal, bit for bit, those of
    a handle that uploaded sigma = 0 and reads it."""
    cfg, mesh, st, eos = make_semi_case(dim, n, bcs=[BoundaryCondition.SlipWall] * 6, seed=11 + dim, tol=1e-10,
                                        maxit=300)
    assert not np.any(st.field(_abi.F_SIGMA))
    runs = []
    for with_sigma in (True, False):
        g = SemiImplicitSolver(mesh, HLLCSolver(eos), JacobiPressureSolver(), eos, None, cfg)
        for fid in st.field_ids():
            if with_sigma or fid != _abi.F_SIGMA:
                g.upload_field(fid, st.field(fid).copy())
        out = []
        for k in range(3):
            cfg.step = k
            out.append((g.step_resident(cfg), g.lastPressureIterations()))
        fields = {f: g.download_field(f).copy() for f in (_abi.F_RHO, _abi.F_RHOU, _abi.F_RHOE, _abi.F_PRES)}
        g.close()
        runs.append((out, fields))
    assert runs[0][0] == runs[1][0]
    for f in runs[0][1]:
        assert np.array_equal(runs[0][1][f], runs[1][1][f]), f
