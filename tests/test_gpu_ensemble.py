"""GPU parity of the batched 1D ensemble (many shock tubes of uneven sizes
in one persistent launch) against one oracle run per tube.

Bar: FP64 — identical step counts and final times, and per tube and field
relative L1 and Linf <= max(1e-12, 3 x the distance between the compiled
reference (ensemble_ref.npz) and the oracle on that tube), against both;
the strongest tubes (near-vacuum rarefactions) put the reference's own FMA
rounding at ~1.5e-12 in velocity and pressure.  Results agree across tile
shapes to rounding (each shape is its own instantiation, and the compiler
may contract multiply-adds differently).  FP32 bulk / FP64 EOS — relative
L1 <= 2e-5 of density per tube against the FP64 oracle.  All calls go
through the C ABI.
"""
import os

import numpy as np
import pytest

from semiimplicitfv_amd import RiemannEnsemble, _abi, cases, ensemble
from tests.golden.problems import ensemble_sample
from tests.helpers import noise_floor_bar, rel_errs

pytestmark = pytest.mark.gpu

T_END = 0.1
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _sample(k=16, nmax=1024, seed=3):
    return ensemble_sample(k, nmax, seed)


def _oracle_tube(oracle, cfg, mesh, st, eos, rs):
    s = oracle.CpuSolver(mesh, cfg, rs.kind(), eos, "c")
    t, n = 0.0, 0
    while t < T_END:
        cfg.step = n
        t += s.explicit_step(st, cfg, T_END - t)
        n += 1
    return n, t


def test_ensemble_fp64_matches_oracle_per_tube(gpu, oracle_c):
    d = np.load(os.path.join(GOLDEN, "ensemble_ref.npz"))
    p, idx = _sample()
    assert list(d["rows"]) == idx
    cfg, eos, rs, tubes = ensemble.build_tubes(p, idx)
    ens = RiemannEnsemble(tubes, cfg, rs, eos)
    steps, times, ms = ens.run(T_END)
    got = {f: ens.download(f) for f in (_abi.F_RHO, _abi.F_RHOU, _abi.F_RHOE, _abi.F_VELU, _abi.F_PRES)}
    ens.close()
    for k, (mesh, st) in enumerate(tubes):
        orc = st.copy()
        n, t = _oracle_tube(oracle_c, cfg, mesh, orc, eos, rs)
        assert steps[k] == n == d[f"t{k}_steps"][0] and times[k] == t == d[f"t{k}_time"][0], (k, steps[k], n)
        for f, arrs in got.items():
            o = mesh.interior(orc.field(f)).ravel()
            r = d[f"t{k}_f{f}"]
            b1, bi = noise_floor_bar(r, o)
            for other in (o, r):
                l1, li = rel_errs(arrs[k], other)
                assert l1 <= b1 and li <= bi, (k, f, l1, li, b1, bi)


def test_ensemble_shapes_bitwise_and_autotune(gpu):
    p, idx = _sample(k=24, nmax=4096, seed=5)
    cfg, eos, rs, tubes = ensemble.build_tubes(p, idx)
    outs = []
    for block, wg in ((64, 4), (256, 2), (1024, 1)):
        ens = RiemannEnsemble(tubes, cfg, rs, eos)
        ens.set_shape(block, wg)
        steps, times, _ = ens.run(T_END, maxSteps=200)
        outs.append((steps, times, ens.download(_abi.F_RHO), ens.download(_abi.F_RHOE)))
        ens.close()
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0])
        assert np.abs(o[1] - outs[0][1]).max() <= 1e-13
        for a, b in zip(o[2] + o[3], outs[0][2] + outs[0][3]):
            l1, li = rel_errs(a, b)
            assert li <= 1e-12, (l1, li)
    # autotune restores the state: a run after tuning equals a fresh run
    ens = RiemannEnsemble(tubes, cfg, rs, eos)
    block, wg, ms = ens.autotune(T_END, maxSteps=20)
    assert block in (64, 128, 256, 512, 1024) and 1 <= wg <= 8 and ms > 0
    ens.set_shape(64, 4)
    steps, times, _ = ens.run(T_END, maxSteps=200)
    assert np.array_equal(steps, outs[0][0]) and np.array_equal(times, outs[0][1])
    for a, b in zip(ens.download(_abi.F_RHO), outs[0][2]):
        assert np.array_equal(a, b)
    ens.close()


def test_ensemble_mixed_precision(gpu, oracle_c):
    p, idx = _sample(k=8, seed=7)
    cfg, eos, rs, tubes = ensemble.build_tubes(p, idx)
    ens = RiemannEnsemble(tubes, cfg, rs, eos, precision="mixed")
    steps, times, _ = ens.run(T_END)
    rho = ens.download(_abi.F_RHO)
    ens.close()
    for k, (mesh, st) in enumerate(tubes):
        ref = st.copy()
        n, t = _oracle_tube(oracle_c, cfg, mesh, ref, eos, rs)
        assert abs(times[k] - T_END) <= 1e-12 and abs(steps[k] - n) <= max(2, n // 100)
        l1, li = rel_errs(rho[k], mesh.interior(ref.rho).ravel())
        assert l1 <= 2e-5, (k, l1, li)


def test_ensemble_rejects_unsupported(gpu):
    cfg, mesh, st, eos, rs = cases.shock_tube(256, (1, 0, 1), (0.125, 0, 0.1))
    cfg.RKOrder = 1
    with pytest.raises(_abi.SifvError):
        RiemannEnsemble([(mesh, st)], cfg, rs, eos)
