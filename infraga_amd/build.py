"""Build recipe for the engine library and the infraga-* front ends (gfx950).

    python -m infraga_amd.build          # library + front ends

Everything is compiled in-tree so the artefacts travel with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libinfraga_gpu.so")
BIN = os.path.join(HERE, "bin")

# -disable-machine-licm: keeps the FP64 polynomial constants of exp / sin /
# Sutherland-Bass from being hoisted into long-lived VGPRs of the step loop.
# -packed-fp32-ops off: required on these boxes (no v_pk_*_f32 generation).
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-mllvm", "-disable-machine-licm",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library():
    src = os.path.join(CSRC, "ga_engine.hip")
    _run(["hipcc", *HIPFLAGS, "-shared", "-o", LIB, src])
    return LIB


def build_frontends():
    os.makedirs(BIN, exist_ok=True)
    src = os.path.join(CSRC, "infraga_cli.cpp")
    if not os.path.exists(src):
        return []
    outs = []
    # infraga-accel-* are the names the reference's Python front end runs
    # under mpirun (cpu_cnt > 1); they take the MPI drivers' launch lists and
    # let rank 0 do the whole job.
    for accel in (0, 1):
        for name, geom in (("infraga-sph", "GA_GEOM_SPH"), ("infraga-3d", "GA_GEOM_3D"),
                           ("infraga-sph-rngdep", "GA_GEOM_SPH_RNGDEP"), ("infraga-2d", "GA_GEOM_2D"),
                           ("infraga-3d-rngdep", "GA_GEOM_3D_RNGDEP")):
            if accel and geom == "GA_GEOM_2D":
                continue
            if accel:
                name = name.replace("infraga-", "infraga-accel-")
            out = os.path.join(BIN, name)
            _run(["g++", "-O2", "-std=c++17", f"-DGA_CLI_GEOM={geom}", f"-DGA_CLI_NAME=\"{name}\"",
                  f"-DGA_CLI_ACCEL={accel}", "-I", os.path.join(ROOT, "include"), "-o", out, src, "-L", HERE,
                  "-linfraga_gpu", f"-Wl,-rpath,$ORIGIN/.."])
            outs.append(out)
    return outs


def build_all():
    build_library()
    build_frontends()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
