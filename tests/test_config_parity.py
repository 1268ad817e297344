"""SimulationConfig.validate mirrors the reference's validate()
(include/SimulationConfig.hpp:123-209): same accept/reject decisions and
messages as the compiled reference on a set of configurations."""
import pytest

from semiimplicitfv_amd import IdealGasEOS, PhaseEOS, ReconstructionOrder, SimulationConfig, uniform_mesh

CASES = [
    ("ok", lambda c: None),
    ("ghost", lambda c: setattr(c, "nGhost", 1)),
    ("rk", lambda c: setattr(c, "RKOrder", 0)),
    ("dim", lambda c: setattr(c, "dim", 4)),
    ("cfl", lambda c: setattr(c.explicitParams, "cfl", -1.0)),
    ("semi_tol", lambda c: (setattr(c, "semiImplicit", True), setattr(c.semiImplicitParams, "pressureTol", 0.0))),
    ("semi_iters", lambda c: (setattr(c, "semiImplicit", True), setattr(c.semiImplicitParams, "maxPressureIters", 0))),
    ("igr", lambda c: (setattr(c, "useIGR", True), setattr(c.igrParams, "IGRIters", 0))),
    ("phase_gamma", lambda c: (setattr(c.multiPhaseParams, "nPhases", 2),
                               setattr(c.multiPhaseParams, "phases", [PhaseEOS(1.0, 0.0), PhaseEOS(1.4, 0.0)]))),
    ("phase_pinf", lambda c: (setattr(c.multiPhaseParams, "nPhases", 2),
                              setattr(c.multiPhaseParams, "phases", [PhaseEOS(4.4, -1.0), PhaseEOS(1.4, 0.0)]))),
]


@pytest.mark.parametrize("name,mut", CASES)
def test_validate_matches_reference(oracle_ref, name, mut):
    cfg = SimulationConfig(dim=1, nGhost=3, RKOrder=3, reconOrder=ReconstructionOrder.WENO5)
    mut(cfg)
    try:
        cfg.validate()
        py = None
    except ValueError as e:
        py = str(e)
    ok = SimulationConfig(dim=1, nGhost=3, RKOrder=3, reconOrder=ReconstructionOrder.WENO5)
    mesh = uniform_mesh(ok, 16, 0.0, 1.0)
    try:
        s = oracle_ref.CpuSolver(mesh, cfg, 2, IdealGasEOS(1.4, 1.0, cfg), "ref")
        ref = None
        s.close()
    except RuntimeError as e:
        ref = str(e).split("ref create failed: ", 1)[1]
    assert py == ref
