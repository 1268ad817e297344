"""Build recipe of libsifv_hip.so (gfx950).

    python -m semiimplicitfv_amd.build

Compiles semiimplicitfv_amd/csrc/*.hip with hipcc for gfx950 into one
shared library next to this file.  Packed FP32 instructions are disabled
(-packed-fp32-ops); hipcc's host pass warns that it does not know that
target feature, which is expected.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsifv_hip.so")
ARCH = os.environ.get("SIFV_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Wno-unused-result", "-Wno-unused-value",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build_torch_extension(verbose: bool = True) -> str:
    """Builds the thin PyTorch extension _sifv_torch (C++, torch API only) in
    place, linked against libsifv_hip.so."""
    import contextlib
    import io

    from setuptools import setup
    from torch.utils.cpp_extension import BuildExtension, CppExtension

    inc = os.path.join(os.path.dirname(HERE), "include")
    ext = CppExtension("semiimplicitfv_amd._sifv_torch", [os.path.join(CSRC, "torch_adjoint.cpp")],
                       include_dirs=[inc], library_dirs=[HERE], libraries=["sifv_hip"],
                       extra_compile_args=["-O2"], extra_link_args=["-Wl,-rpath,$ORIGIN"])
    t0 = time.time()
    cwd = os.getcwd()
    os.chdir(os.path.dirname(HERE))
    try:
        out = io.StringIO()
        with contextlib.redirect_stdout(out), contextlib.redirect_stderr(out):
            setup(name="sifv_torch", ext_modules=[ext], cmdclass={"build_ext": BuildExtension},
                  script_args=["-q", "build_ext", "--inplace", "--build-temp", os.path.join(HERE, "_build"),
                               "--build-lib", os.path.join(HERE, "_build", "lib")])
    finally:
        os.chdir(cwd)
    if verbose:
        print(f"built the torch extension in {time.time() - t0:.1f} s")
    import glob as _g
    return (_g.glob(os.path.join(HERE, "_sifv_torch*.so")) or [""])[0]


def build(verbose: bool = True) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = []
    t0 = time.time()
    procs = []
    for s in srcs:
        o = os.path.join(CSRC, os.path.basename(s) + ".o")
        cmd = [hipcc] + [f for f in FLAGS if f != "-shared"] + ["-c", s, "-o", o]
        procs.append((s, o, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    for s, o, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode(errors="replace"))
            raise RuntimeError(f"hipcc failed on {s}")
        objs.append(o)
    tmp = OUT + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-soname,libsifv_hip.so", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode(errors="replace"))
        raise RuntimeError("link of libsifv_hip.so failed")
    os.replace(tmp, OUT)
    for o in objs:
        os.remove(o)
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources in {time.time() - t0:.1f} s")
    return OUT


if __name__ == "__main__":
    build()
    build_torch_extension()
