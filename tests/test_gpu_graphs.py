"""Graph replay of the explicit RK step (sifv_explicit_advance): the same
launches replayed from a captured HIP graph give bit-identical states, time
and step counts to enqueuing every launch, across several advance calls with
host work in between, an end time reached mid-call, RK2/RK3, 2D/3D and the
FP32-bulk form.  All calls go through the C ABI."""
import numpy as np
import pytest

from semiimplicitfv_amd import BoundaryCondition, ExplicitSolver, HLLCSolver, RusanovSolver, cases
from tests.golden.problems import make_case

pytestmark = pytest.mark.gpu


def _vortex(precision):
    cfg, mesh, st, eos, rs = cases.isentropic_vortex_2d(64)
    return cfg, mesh, st, eos, rs, precision


def _case3d(precision):
    cfg, mesh, st, eos = make_case(3, (20, 18, 16), bcs=[BoundaryCondition.Outflow] * 6, seed=3, nonuniform=True)
    return cfg, mesh, st, eos, HLLCSolver(eos), precision


def _rk2(precision):
    cfg, mesh, st, eos = make_case(2, (48, 40), periods=(1, 1, 0), rk=2, seed=8)
    return cfg, mesh, st, eos, RusanovSolver(eos), precision


CASES = {"vortex2d_rk3": lambda: _vortex("fp64"), "vortex2d_mixed": lambda: _vortex("mixed"),
         "case3d_rk3": lambda: _case3d("fp64"), "periodic2d_rk2": lambda: _rk2("fp64")}


@pytest.mark.parametrize("name", sorted(CASES))
def test_graph_replay_matches_launches(gpu, name):
    cfg, mesh, st, eos, rs, precision = CASES[name]()
    solvers = []
    for graphs in (True, False):
        g = ExplicitSolver(mesh, rs, eos, None, cfg, precision=precision)
        g.set_graphs(graphs)
        g.upload(st)
        solvers.append(g)
    # a first probe fixes an end time that the third call crosses
    probe = ExplicitSolver(mesh, rs, eos, None, cfg, precision=precision)
    probe.upload(st)
    t_probe, _ = probe.advance(cfg, 25, 1e30, 0.0)
    probe.close()
    t_end = t_probe
    hist = [[], []]
    for k, nsteps in enumerate((7, 13, 9, 4)):
        for q, g in enumerate(solvers):
            t0 = hist[q][-1][0] if hist[q] else 0.0
            t, n = g.advance(cfg, nsteps, t_end, t0)
            hist[q].append((t, n))
            # host work between calls: a download of one field
            rho = g.download_field(0)
            assert np.isfinite(mesh.interior(rho)).all()
    assert hist[0] == hist[1], hist
    for fid in st.field_ids():
        a, b = solvers[0].download_field(fid), solvers[1].download_field(fid)
        assert np.array_equal(mesh.interior(a), mesh.interior(b)), fid
    assert solvers[0].graph_replays() > 0
    assert solvers[1].graph_replays() == 0
    for g in solvers:
        g.close()


def test_graphs_default_off(gpu):
    cfg, mesh, st, eos, rs = cases.isentropic_vortex_2d(32)
    g = ExplicitSolver(mesh, rs, eos, None, cfg)
    g.upload(st)
    g.advance(cfg, 3, 1e30, 0.0)
    assert g.graph_replays() == 0
    g.close()


def test_no_graph_while_timing(gpu):
    """Per-launch timing records events around each launch: no graph then."""
    cfg, mesh, st, eos, rs = cases.isentropic_vortex_2d(32)
    g = ExplicitSolver(mesh, rs, eos, None, cfg)
    g.set_graphs(True)
    g.upload(st)
    g.advance(cfg, 3, 1e30, 0.0)
    n0 = g.graph_replays()
    assert n0 > 0
    g.timing(True)
    g.advance(cfg, 4, 1e30, 0.0)
    ms, launches = g.read_timing(reset=True)
    g.timing(False)
    assert g.graph_replays() == n0 and launches == 12 and ms > 0.0
    g.close()
