"""Build libcosims_engine.so (HIP kernels for gfx950 + host setup + C ABI)
and the command-line program that drives it.

Compiles in-tree: objects under cosims_amd/_build/, the library at
cosims_amd/libcosims_engine.so, the program at cosims_amd/cosims, the VALU
issue-rate probes at cosims_amd/issue_probe.  Run
``python -m cosims_amd.build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcosims_engine.so")
CLI = os.path.join(HERE, "cosims")
PROBE = os.path.join(HERE, "issue_probe")
ARCH = os.environ.get("COSIMS_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"

# -ffp-contract=off: the FP64 path keeps the reference's unfused IEEE
# arithmetic; the FP32 evaluators opt back into contraction locally.
# -packed-fp32-ops: no v_pk_*_f32 generation.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wno-unused-result"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-variable"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + SRC]
    objs = []
    for name in ("prep.cpp", "io.cpp"):
        o = os.path.join(BUILD, name + ".o")
        _run([CXX, *CXX_FLAGS, *inc, "-c", os.path.join(SRC, name), "-o", o], verbose)
        objs.append(o)
    o = os.path.join(BUILD, "engine.o")
    _run([HIPCC, *HIP_FLAGS, *inc, "-c", os.path.join(SRC, "engine.hip"), "-o", o], verbose)
    objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"], verbose)
    os.replace(tmp, LIB)
    # the command-line program: host code over the C ABI, found next to the library
    o = os.path.join(BUILD, "cli.o")
    _run([CXX, *CXX_FLAGS, *inc, "-c", os.path.join(SRC, "cli.cpp"), "-o", o], verbose)
    tmp = CLI + ".tmp"
    _run([HIPCC, "-o", tmp, o, "-L" + HERE, "-lcosims_engine", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, CLI)
    # the VALU issue-rate probes behind the roofline figures (a measurement program)
    tmp = PROBE + ".tmp"
    _run([HIPCC, *HIP_FLAGS, *inc, "-o", tmp, os.path.join(SRC, "issue_probe.hip")], verbose)
    os.replace(tmp, PROBE)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
