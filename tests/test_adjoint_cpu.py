"""Pins of the tangent-mode restatement (oracle/tangent.py) that checks the
GPU discrete adjoint: its primal part against the C oracle, its tangent
against central differences of its own primal part."""
import numpy as np
import pytest

from oracle import tangent
from semiimplicitfv_amd import cases
from tests.helpers import rel_errs


def _vortex(n):
    """Perturbed vortex (no mirror symmetry, so no exact ties in the
    Rusanov wave-speed max, where the discrete map has a kink)."""
    cfg, mesh, st, eos, rs, _ = cases.perturbed_vortex_2d(n, amplitude=0.05, width=2.0)
    h = 10.0 / n
    return cfg, mesh, st, eos, rs, cfg.explicitParams.constDt, (h, h)


def _interior(mesh, st, dim):
    ids = [0, 1, 2, 4] if dim == 2 else [0, 1, 4]
    return [np.ascontiguousarray(mesh.interior(st.field(f))[0]) for f in ids]


def test_tangent_primal_matches_oracle(oracle_c):
    cfg, mesh, st, eos, rs, dt, dx = _vortex(24)
    U0 = _interior(mesh, st, 2)
    s = oracle_c.CpuSolver(mesh, cfg, rs.kind(), eos, "c")
    for k in range(4):
        cfg.step = k
        assert s.explicit_step(st, cfg) == dt
    U, _ = tangent.run(U0, [np.zeros_like(a) for a in U0], 1.4, 0.0, dt, 4, dx)
    ref = _interior(mesh, st, 2)
    for a, b in zip(U, ref):
        l1, li = rel_errs(a.v, b)
        assert l1 <= 1e-13 and li <= 1e-12, (l1, li)


@pytest.mark.parametrize("seed", [0, 1])
def test_tangent_matches_central_differences(seed):
    cfg, mesh, st, eos, rs, dt, dx = _vortex(16)
    U0 = _interior(mesh, st, 2)
    rng = np.random.default_rng(seed)
    d = [rng.standard_normal(a.shape) * (1e-2 * np.abs(a).max()) for a in U0]
    dg = 0.05 * rng.standard_normal()
    _, J = tangent.run(U0, d, 1.4, dg, dt, 3, dx)
    # the discrete map has kinks (the Rusanov wave-speed max), so central
    # differences converge only to ~1e-7 here; h = 1e-6 sits at the optimum
    h = 1e-6
    zero = [np.zeros_like(a) for a in U0]
    _, Jp = tangent.run([a + h * b for a, b in zip(U0, d)], zero, 1.4 + h * dg, 0.0, dt, 3, dx)
    _, Jm = tangent.run([a - h * b for a, b in zip(U0, d)], zero, 1.4 - h * dg, 0.0, dt, 3, dx)
    fd = (Jp.v - Jm.v) / (2 * h)
    assert abs(J.t - fd) <= 2e-6 * abs(fd), (J.t, fd)
