"""Build recipe for libpepsmg.so (HIP, gfx950) and the Fortran drop-in.

Run ``python -m peps_mg_amd.build`` (or __graft_entry__.build()).  The
library is built in-tree next to this file so it travels with the source
tree; nothing is installed.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpepsmg.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PEPSMG_ARCH", "gfx950")

SOURCES = ["pepsmg_core.hip", "pepsmg_solve.hip", "pepsmg_batch.hip"]

# -ffp-contract=off: no fused multiply-add anywhere, so FP64 results are
# the reference's bit for bit.  -packed-fp32-ops: no packed FP32 VALU ops.
HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function", "-Wno-unused-value",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_library(verbose=True):
    """Compiles the translation units in parallel (one hipcc each), then links."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [HIPCC] + compile_flags + ["-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), cmd))
        objs.append(obj)
    failed = [cmd for p, cmd in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-fPIC", "-shared"] + objs + ["-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_all():
    build_library()
    from peps_mg_amd import fortran_build
    fortran_build.build()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "lib"
    if what == "all":
        build_all()
    else:
        build_library()
