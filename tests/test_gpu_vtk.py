"""VTK output of the resident state (sifv_write_vtr / _pvtr / _pvd) against
the compiled reference's VTKWriter (oracle/_ref), and the resident output
loop (semiimplicitfv_amd.vtk.run_time_loop_vtk) against the reference's
runTimeLoop + VTKSession on the same problem.

1. Byte parity: the same state (seeded fields, 1D single-phase, 2D two-phase,
   3D with IGR sigma) uploaded to a GPU handle and written by the library is
   the reference's file byte for byte, for the rank-0 piece, the serial form,
   the .pvtr and a .pvd sequence.
2. Resident run: Sod-256 to t = 0.2 with output every 0.05 writes the same
   files as the reference run; the numbers agree to the FP64 parity bar
   (the two runs differ by rounding, so the 15th printed digit may differ).
"""
import os
import re

import numpy as np
import pytest

from semiimplicitfv_amd import ExplicitSolver, HLLCSolver, LFSolver, cases
from semiimplicitfv_amd.timeloop import run_time_loop
from semiimplicitfv_amd.vtk import VTKSession, run_time_loop_vtk
from tests.golden.problems import make_case, make_two_phase_case

pytestmark = pytest.mark.gpu


def _seeded(cfg, mesh, st, seed):
    rng = np.random.default_rng(seed)
    for fid in st.field_ids():
        a = st.field(fid)
        a[:] = rng.standard_normal(a.size) * 10.0 ** rng.integers(-3, 4)
    return st


def _states():
    cfg1, m1, s1, e1, rs1 = cases.sod_1d(37)
    cfg2, m2, s2, e2 = make_two_phase_case(2, (13, 9), seed=3)
    cfg3, m3, s3, e3 = make_case(3, (7, 6, 5), seed=4, nonuniform=True)
    cfg3.useIGR = True
    cfg3.validate()
    s3 = s3.copy()
    return {"sod1d": (cfg1, m1, _seeded(cfg1, m1, s1, 1), e1, rs1),
            "twophase2d": (cfg2, m2, _seeded(cfg2, m2, s2, 2), e2, HLLCSolver(e2)),
            "igr3d": (cfg3, m3, _seeded(cfg3, m3, s3, 3), e3, LFSolver(e3))}


@pytest.mark.parametrize("name", ["sod1d", "twophase2d", "igr3d"])
def test_vtk_bytes_match_reference(gpu, oracle_ref, tmp_path, name):
    cfg, mesh, st, eos, rs = _states()[name]
    g = ExplicitSolver(mesh, rs, eos, None, cfg)
    g.upload(st)
    ref = oracle_ref.CpuSolver(mesh, cfg, rs.kind(), eos, "ref")
    lib = g._lib
    for rank in (0, -1):
        a, b = tmp_path / f"gpu_{rank}.vtr", tmp_path / f"ref_{rank}.vtr"
        assert lib.sifv_write_vtr(g._h, str(a).encode(), rank) == 0
        ref.write_vtr(st, str(b), rank)
        assert a.read_bytes() == b.read_bytes(), (name, rank)
    a, b = tmp_path / "gpu.pvtr", tmp_path / "ref.pvtr"
    assert lib.sifv_write_pvtr(g._h, str(a).encode(), b"piece_0_r0.vtr") == 0
    ref.write_pvtr(str(b), "piece_0_r0.vtr")
    assert a.read_bytes() == b.read_bytes()
    a, b = tmp_path / "sub" / "gpu.pvd", tmp_path / "sub" / "ref.pvd"
    for mode, t, f in (("w", 0.0, ""), ("a", 0.0, "s0/x.pvtr"), ("a", 0.123456789012345678, "s1/x.pvtr"),
                       ("a", 1e-7 / 3.0, "s2/x.pvtr"), ("close", 0.0, "")):
        assert lib.sifv_write_pvd(str(a).encode(), mode.encode(), t, f.encode()) == 0
        ref.write_pvd(str(b), mode, t, f)
    assert a.read_bytes() == b.read_bytes()
    g.close()


class _RefSession:
    """VTKSession (src/VTKSession.cpp) over the compiled reference's writer."""

    def __init__(self, ref, st, base, d):
        self.ref, self.st, self.base, self.dir, self.n = ref, st, base, d, 0
        ref.write_pvd(os.path.join(d, base + ".pvd"), "w")

    def write(self, time):
        snap, vtr, pvtr = f"snapshot_{self.n}", f"{self.base}_{self.n}_r0.vtr", f"{self.base}_{self.n}.pvtr"
        self.ref.write_vtr(self.st, os.path.join(self.dir, snap, vtr), 0)
        self.ref.write_pvtr(os.path.join(self.dir, snap, pvtr), vtr)
        self.ref.write_pvd(os.path.join(self.dir, self.base + ".pvd"), "a", time, f"{snap}/{pvtr}")
        self.n += 1


_NUM = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?|nan|inf")


def _tokens(text):
    nums = [float(x) for x in _NUM.findall(text)]
    return _NUM.sub("#", text), np.array(nums)


def test_resident_output_loop_matches_reference(gpu, oracle_ref, tmp_path):
    cfg, mesh, st, eos, rs = cases.sod_1d(256)
    T, dT = 0.2, 0.05
    gdir, rdir = tmp_path / "gpu", tmp_path / "ref"
    g = ExplicitSolver(mesh, rs, eos, None, cfg)
    g.upload(st)
    vtk = VTKSession(g, "sod", str(gdir))
    tg, ng = run_time_loop_vtk(cfg, g, vtk, lambda tdt: g.step_resident(cfg, tdt), T, dT)
    g.close()

    cfg2, mesh2, st2, eos2, rs2 = cases.sod_1d(256)
    ref = oracle_ref.CpuSolver(mesh2, cfg2, rs2.kind(), eos2, "ref")
    rsess = _RefSession(ref, st2, "sod", str(rdir))

    def out(t):
        ref.cons_to_prim(st2, cfg2)
        rsess.write(t)

    ref.cons_to_prim(st2, cfg2)
    rsess.write(0.0)
    tr, nr = run_time_loop(cfg2, lambda tdt: ref.explicit_step(st2, cfg2, tdt), T, dT, on_output=out)
    assert ng == nr and abs(tg - tr) <= 1e-14
    files_g = sorted(p.relative_to(gdir) for p in gdir.rglob("*") if p.is_file())
    files_r = sorted(p.relative_to(rdir) for p in rdir.rglob("*") if p.is_file())
    assert files_g == files_r and len(files_g) == 1 + 2 * 5
    for rel in files_g:
        tg_, ng_ = _tokens((gdir / rel).read_text())
        tr_, nr_ = _tokens((rdir / rel).read_text())
        assert tg_ == tr_, rel
        assert ng_.shape == nr_.shape
        scale = np.maximum(1.0, np.abs(nr_))
        assert np.all(np.abs(ng_ - nr_) <= 1e-12 * scale), (rel, np.abs(ng_ - nr_).max())
