"""The C++ drop-in (include/sifv_gpu_solvers.hpp) compiles against the
reference's headers and links with the reference's sources (built by
oracle/Makefile into oracle/_ref/dropin_driver); on the GPU the reference's
CPU ExplicitSolver and gpu::ExplicitSolver run Sod-1024 side by side."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_driver")


def test_dropin_driver_was_built():
    if not os.path.isdir("/root/reference"):
        pytest.skip("reference checkout absent")
    assert os.path.exists(EXE), "oracle/Makefile did not build the C++ drop-in driver"


@pytest.mark.gpu
def test_dropin_driver_runs_on_gpu():
    if not os.path.exists(EXE):
        pytest.skip("drop-in driver not built (needs the reference checkout at build time)")
    r = subprocess.run([EXE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=240)
    print(r.stdout.decode())
    assert r.returncode == 0, r.stdout.decode()
