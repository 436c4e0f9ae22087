"""C-ABI checks that need no GPU: the library loads and exports every
function include/pepsmg.h declares, and the ctypes table matches."""
import os
import re

import pytest

from peps_mg_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "pepsmg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pepsmg_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_expected_entry_points():
    names = header_functions()
    for n in ["pepsmg_multigrid_create", "pepsmg_multigrid_solve_vcycle", "pepsmg_multigrid_destroy",
              "pepsmg_rbgs_iterator", "pepsmg_rbgs_solver", "pepsmg_cg_iterator", "pepsmg_cg_solver",
              "pepsmg_mv_mul", "pepsmg_vv_dot"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for n in header_functions():
        assert hasattr(L, n), f"{n} not exported"


def test_ctypes_table_covers_header():
    assert sorted(_lib.SIGNATURES) == header_functions()


def test_default_options_without_gpu():
    o = _lib.default_options()
    assert o.smoother == 0 and o.precision == 0 and o.coarse_switch == 4096 and o.fused_smoother == 1


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(RuntimeError, match="missing"):
        _lib.lib()


def test_struct_layouts_match_header(tmp_path):
    """sizeof and every field offset of the three structs, as a C compiler
    lays them out from include/pepsmg.h, equal the ctypes mirrors'."""
    import ctypes as C
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    mirrors = {"pepsmg_grid": _lib.Grid, "pepsmg_options": _lib.Options, "pepsmg_stats": _lib.Stats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pepsmg.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


@pytest.mark.parametrize("ax,nx", [(1.1, 15), (1.0, 16)])
def test_degenerate_grid_rejected_without_gpu(ax, nx):
    """The grid check runs before any HIP call: a stretched axis with an odd
    cell count (src/geometry.f90:190-197 leaves its middle width unset) and a
    zero width written into a uniform axis are both PEPSMG_ERR_ARG."""
    from peps_mg_amd import geometry
    from peps_mg_amd.multigrid import Multigrid
    sdm = geometry.geometry_subdomain_create(
        geometry.geometry_domain_create(nx, 8, 8, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, ax=ax))
    if ax == 1.0:
        sdm.dxm[5] = 0.0
    with pytest.raises(_lib.PepsmgError) as e:
        Multigrid(sdm, 2, 10)
    assert e.value.code == _lib.ERR_ARG
    assert "x axis" in str(e.value)
