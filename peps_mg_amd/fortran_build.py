"""Build recipe for the Fortran drop-in (amdflang).

Compiles the modules that keep the reference's module and procedure names
(mpi_topology, geometry, matrix, poisson_matrix_operator,
rbgs_poisson_matrix, cg_poisson_matrix, multigrid) over the iso_c_binding
interface to libpepsmg.so, into peps_mg_amd/fortran/build/: the module
files, a static library libpepsmg_f.a for callers that bring their own main
program, and the driver main_peps (reads PARA_INPUT.inp from the working
directory).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FSRC = os.path.join(HERE, "fortran")
BUILD = os.path.join(FSRC, "build")
FC = os.environ.get("PEPSMG_FC", "/opt/rocm/bin/amdflang")

MODULES = ["pepsmg_c", "mpi_topology", "geometry", "matrix", "pepsmg_shim", "poisson_matrix_operator",
           "rbgs_poisson_matrix", "cg_poisson_matrix", "multigrid"]
FFLAGS = ["-O2"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=BUILD)


def build():
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for m in MODULES:
        src = os.path.join(FSRC, m + ".f90")
        obj = os.path.join(BUILD, m + ".o")
        _run([FC] + FFLAGS + ["-module-dir", BUILD, "-c", src, "-o", obj])
        objs.append(obj)
    lib = os.path.join(BUILD, "libpepsmg_f.a")
    if os.path.exists(lib):
        os.remove(lib)
    _run(["ar", "rcs", lib] + objs)
    exe = os.path.join(BUILD, "main_peps")
    _run([FC] + FFLAGS + ["-module-dir", BUILD, os.path.join(FSRC, "main_peps.f90"), lib,
                          "-L" + HERE, "-lpepsmg", "-Wl,-rpath," + HERE, "-o", exe])
    return exe


if __name__ == "__main__":
    build()
    sys.exit(0)
