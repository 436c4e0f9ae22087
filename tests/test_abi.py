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
