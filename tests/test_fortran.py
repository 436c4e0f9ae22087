"""Fortran drop-in (peps_mg_amd/fortran): build, CPU-side behaviour, and the
driver's solution against the oracle on the GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

from peps_mg_amd import fortran_build

EXE = os.path.join(fortran_build.BUILD, "main_peps")

PARA = """&meshes
nx = {n}
ny = {n}
nz = {n}
/
&origin
ox = {o}
oy = {o}
oz = {o}
/
&length
lx = 1.0d0
ly = 1.0d0
lz = {lz}
/
&mesh_stretch
ax = 1.0d0
ay = 1.0d0
az = 1.0d0
/
&procs
npx = {npx}
npy = 1
npz = 1
/
&control
maxiteration = 2
tolerance = 1.0d-10
number_of_vcycles = 60
number_of_levels = {nl}
aggregation_method = {aggr}
CGA_level = 0
CGPSA_level = 0, 0, 0
omega_sor = 1.0d0
/
&coefficients
alpha_x = 1.0d0
alpha_y = 1.0d0
alpha_z = 1.0d0
/
"""


def ensure_built():
    if not os.path.exists(EXE):
        fortran_build.build()
    return EXE


def write_input(d, n=16, lz="1.0d0", npx=1, nl=2, aggr=0, gpu=None, o="0.0d0"):
    txt = PARA.format(n=n, lz=lz, npx=npx, nl=nl, aggr=aggr, o=o)
    if gpu:
        txt += "&gpu\n" + "\n".join(f"{k} = {v}" for k, v in gpu.items()) + "\n/\n"
    with open(os.path.join(d, "PARA_INPUT.inp"), "w") as f:
        f.write(txt)


def test_fortran_dropin_builds_with_reference_module_names():
    ensure_built()
    for m in ["mpi_topology", "geometry", "matrix", "poisson_matrix_operator", "rbgs_poisson_matrix",
              "cg_poisson_matrix", "multigrid"]:
        assert os.path.exists(os.path.join(fortran_build.BUILD, m + ".mod")), m
    syms = subprocess.run(["nm", os.path.join(fortran_build.BUILD, "libpepsmg_f.a")], capture_output=True,
                          text=True, check=True).stdout
    for p in ["multigrid_create", "multigrid_solve_vcycle", "multigrid_destroy", "rbgs_iterator_poisson_matrix",
              "rbgs_solver_poisson_matrix", "cg_iterator_poisson_matrix", "cg_solver_poisson_matrix",
              "mv_mul_poisson_matrix", "vv_dot_3d_matrix", "matrix_heptadiagonal_create", "geometry_domain_create"]:
        assert re.search(p, syms, re.I), p


def test_driver_rejects_several_processes(tmp_path):
    """npx > 1 stops before any GPU call (one GPU owns the grid)."""
    exe = ensure_built()
    write_input(tmp_path, npx=4)
    p = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert p.returncode != 0
    assert "one process" in p.stdout


@pytest.mark.gpu
def test_driver_matches_oracle(tmp_path):
    from oracle import oracle as O
    from peps_mg_amd import _lib
    exe = ensure_built()
    n, nl = 32, 3
    write_input(tmp_path, n=n, nl=nl, gpu=dict(rhs_mode=1, seed=7, output_file="'x.bin'"))
    p = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    cyc = [float(m.group(1)) for m in re.finditer(r"\[MG\] cycle =\s+\d+, Error:\s+\S+\s+\S+\s+(\S+)", p.stdout)]
    x = np.fromfile(tmp_path / "x.bin", dtype=np.float64).reshape(n + 2, n + 2, n + 2)
    g = O.Grid(n, n, n)
    b = O.padded(n, n, n)
    _lib.check(_lib.lib().pepsmg_fill_rhs_host(n, n, n, _lib.dptr(b), 7, -1.0, -3 * np.pi ** 2, _lib.dptr(g.xg),
                                               _lib.dptr(g.yg), _lib.dptr(g.zg)))
    mo = O.Multigrid(g, nl, 60)
    xo, ro = O.padded(n, n, n), O.padded(n, n, n)
    nc, hist, cit, res0 = mo.solve_vcycle(xo, ro, g.matrix(), b, 2, 1e-10, 1.0)
    assert len(cyc) == nc
    np.testing.assert_allclose(cyc, hist, rtol=1e-6)       # printed with 8 significant digits
    assert np.array_equal(x, xo)


@pytest.mark.gpu
def test_driver_mixed_precision_inner_schedule(tmp_path):
    """precision = 1 with an asymmetric inner V-cycle set through the &gpu
    namelist: the driver's solution meets the tolerance against an
    independent FP64 residual and stays within 1e-10 of the oracle's FP64
    solution (the workload's right-hand side: a 1e-10 residual puts either
    solution within 2e-11 of the exact one on this grid)."""
    import refmath as RM
    from oracle import oracle as O
    from peps_mg_amd import _lib
    exe = ensure_built()
    n, nl = 32, 3
    write_input(tmp_path, n=n, nl=nl, gpu=dict(rhs_mode=1, seed=5, output_file="'x.bin'", precision=1,
                                               inner_nu=2, inner_nu_post=1, inner_omega="1.2d0"))
    p = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    x = np.fromfile(tmp_path / "x.bin", dtype=np.float64).reshape(n + 2, n + 2, n + 2)
    g = O.Grid(n, n, n)
    b = O.padded(n, n, n)
    _lib.check(_lib.lib().pepsmg_fill_rhs_host(n, n, n, _lib.dptr(b), 5, -1.0, -3 * np.pi ** 2, _lib.dptr(g.xg),
                                               _lib.dptr(g.yg), _lib.dptr(g.zg)))
    K = RM.laplacian(g.dxm[1:-1], g.dym[1:-1], g.dzm[1:-1])
    r = RM.flat(b) - K @ RM.flat(x)
    assert np.linalg.norm(r) / np.linalg.norm(RM.flat(b)) < 1.01e-10
    mo = O.Multigrid(g, nl, 60)
    xo, ro = O.padded(n, n, n), O.padded(n, n, n)
    mo.solve_vcycle(xo, ro, g.matrix(), b, 2, 1e-10, 1.0)
    assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()


@pytest.mark.gpu
def test_driver_cosine_problem(tmp_path):
    """The reference's Problem 1 (cosine RHS over -l/2 <= x,y,z <= l/2)
    converges with O(h^2) RMS error."""
    exe = ensure_built()
    write_input(tmp_path, n=32, nl=3, o="-0.5d0")
    p = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    rms = float(re.search(r"RMS =\s+(\S+)", p.stdout).group(1))
    assert rms < 1e-5, p.stdout
    assert "converged =   1" in p.stdout


@pytest.mark.gpu
def test_driver_level_error_stops(tmp_path):
    exe = ensure_built()
    write_input(tmp_path, n=16, nl=9)
    p = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "multigrid_create failed" in p.stdout


@pytest.mark.gpu
def test_driver_writes_reference_solution_files(tmp_path):
    """write_solution: the reference's solution_all.dat / .bin / _MPI_IO.bin
    (src/main_peps.f90:204-446) hold the interior of the solved field."""
    exe = ensure_built()
    n = 12
    write_input(tmp_path, n=n, nl=2, gpu=dict(rhs_mode=1, seed=3, output_file="'x.bin'", write_solution=".true."))
    p = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "Solution printed" in p.stdout
    x = np.fromfile(tmp_path / "x.bin", dtype=np.float64).reshape(n + 2, n + 2, n + 2)[1:-1, 1:-1, 1:-1]
    raw = np.fromfile(tmp_path / "solution_all_MPI_IO.bin", dtype=np.float64).reshape(n, n, n)
    assert np.array_equal(raw, x)
    rec = (tmp_path / "solution_all.bin").read_bytes()
    size = np.frombuffer(rec[:4], dtype=np.int32)[0]
    assert size == n ** 3 * 8 and np.frombuffer(rec[-4:], dtype=np.int32)[0] == size
    assert np.array_equal(np.frombuffer(rec[4:-4], dtype=np.float64).reshape(n, n, n), x)
    lines = (tmp_path / "solution_all.dat").read_text().splitlines()
    assert len(lines) == n ** 3
    first = [float(v.replace("D", "E")) for v in lines[0].split()]
    assert abs(first[3] - x[0, 0, 0]) <= 1e-18 + 1e-19 * abs(x[0, 0, 0])
    last = [float(v.replace("D", "E")) for v in lines[-1].split()]
    assert abs(last[3] - x[-1, -1, -1]) <= 1e-18 + 1e-19 * abs(x[-1, -1, -1])
    assert len(lines[0]) == 3 * 20 + 30
    # the gatherv and gatherv_halo collections (src/main_peps.f90:321-400)
    # hold the same interior on one process
    for name in ("solution_all_gatherv", "solution_all_gatherv_halo"):
        assert (tmp_path / (name + ".bin")).read_bytes() == rec, name
        assert (tmp_path / (name + ".dat")).read_text().splitlines() == lines, name


def test_bind_c_types_match_the_c_structs(tmp_path):
    """The Fortran bind(C) mirrors of pepsmg_grid, pepsmg_options and
    pepsmg_stats have the sizes of the C structs (whose field offsets
    test_abi checks against a C compiler)."""
    import ctypes as C
    from peps_mg_amd import _lib
    if not os.path.exists(fortran_build.FC):
        pytest.skip("no Fortran compiler")
    ensure_built()
    src = tmp_path / "layout.f90"
    src.write_text("program layout\n"
                   "    use iso_c_binding\n"
                   "    use pepsmg_c\n"
                   "    type(pepsmg_grid_t) :: g\n"
                   "    type(pepsmg_options_t) :: o\n"
                   "    type(pepsmg_stats_t) :: s\n"
                   "    print '(3i8)', c_sizeof(g), c_sizeof(o), c_sizeof(s)\n"
                   "end program layout\n")
    exe = tmp_path / "layout"
    subprocess.run([fortran_build.FC, "-I", fortran_build.BUILD, str(src), "-o", str(exe)], check=True,
                   cwd=tmp_path)
    sizes = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                            text=True).stdout.split()]
    assert sizes == [C.sizeof(_lib.Grid), C.sizeof(_lib.Options), C.sizeof(_lib.Stats)]


SAMPLE = """&meshes
nx = 1024
ny = 64
nz = 512
/
&origin
ox = -8.0d0
oy = -0.5d0
oz = -4.0d0
/
&length
lx = 16.0d0
ly = 1.0d0
lz = 8.0d0
/
&mesh_stretch
ax = 1.0d0
ay = 1.0d0
az = 1.0d0
/
&procs
npx = 1
npy = 1
npz = 1
/
&control
maxiteration = 1
tolerance = 1.0d-10
number_of_vcycles = 3
number_of_levels = 10
aggregation_method = 2
CGA_level = 6
CGPSA_level = 9, 6, 8
omega_sor = 1.6d0
/
&coefficients
alpha_x = 1.0d0
alpha_y = 1.0d0
alpha_z = 1.0d0
/
&gpu
output_file = 'x.bin'
rhs_file = 'b.bin'
/
"""


@pytest.mark.gpu
def test_driver_reference_sample_input(tmp_path):
    """The reference's sample input (run/PARA_INPUT.inp) on one process with
    every edge halved (1024 x 64 x 512 cells on the same box, so the same
    h = 1/64 spacing: number_of_levels 10, the halvings x allows under the
    CGPSA rule of src/multigrid.f90:1186-1193): Problem 1's cosine
    right-hand side, V(1,1), omega 1.6, stopped by the cycle cap after 3
    cycles.  The printed [MG] history (8 digits) and the solution (bit for
    bit) equal the oracle's from the same right-hand side."""
    from oracle import oracle as O
    exe = ensure_built()
    (tmp_path / "PARA_INPUT.inp").write_text(SAMPLE)
    p = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "Not converged" in p.stdout
    nx, ny, nz = 1024, 64, 512
    cyc = [float(m.group(1)) for m in re.finditer(r"\[MG\] cycle =\s+\d+, Error:\s+\S+\s+\S+\s+(\S+)", p.stdout)]
    b = np.fromfile(tmp_path / "b.bin", dtype=np.float64).reshape(nz + 2, ny + 2, nx + 2)
    x = np.fromfile(tmp_path / "x.bin", dtype=np.float64).reshape(nz + 2, ny + 2, nx + 2)
    g = O.Grid(nx, ny, nz, o=(-8.0, -0.5, -4.0), l=(16.0, 1.0, 8.0))
    mo = O.Multigrid(g, 10, 3, 2, 6)
    xo, ro = O.padded(nx, ny, nz), O.padded(nx, ny, nz)
    nc, hist, cit, res0 = mo.solve_vcycle(xo, ro, g.matrix(), b, 1, 1e-10, 1.6)
    assert nc == 3 and len(cyc) == 3
    np.testing.assert_allclose(cyc, hist, rtol=1e-6)       # printed with 8 significant digits
    assert np.array_equal(x, xo)


def test_fortran_extent_checks_stop_before_the_library(tmp_path):
    """An array whose extents do not match the grid stops the call with an
    [Error] line before anything is copied (no GPU involved)."""
    if not os.path.exists(fortran_build.FC):
        pytest.skip("no Fortran compiler")
    ensure_built()
    src = tmp_path / "extents.f90"
    src.write_text("program extents\n"
                   "    use poisson_matrix_operator\n"
                   "    real(kind=8), allocatable :: x(:,:,:), y(:,:,:)\n"
                   "    real(kind=8) :: r\n"
                   "    allocate(x(0:9,0:9,0:9), y(0:9,0:9,0:7))\n"
                   "    x = 0.0d0\n"
                   "    y = 0.0d0\n"
                   "    call vv_dot_3d_matrix(r, x, y, 8, 8, 8, [.true., .true., .true.])\n"
                   "    print *, 'not reached', r\n"
                   "end program extents\n")
    from peps_mg_amd import _lib
    exe = tmp_path / "extents"
    subprocess.run([fortran_build.FC, "-I", fortran_build.BUILD, str(src),
                    os.path.join(fortran_build.BUILD, "libpepsmg_f.a"), "-L" + _lib.HERE, "-lpepsmg",
                    "-Wl,-rpath," + _lib.HERE, "-o", str(exe)], check=True, cwd=tmp_path)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0
    assert "[Error] vv_dot_3d_matrix: y has extents" in p.stdout and "not reached" not in p.stdout
