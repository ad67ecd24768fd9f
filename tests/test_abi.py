"""The C-ABI library loads without a GPU and exports every entry point that
include/nise_abi.h declares (no compute calls here)."""
import ctypes as C
import os
import re

from nise_2017_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "nise_abi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_header_declares_the_drop_in_symbols():
    names = declared_functions()
    for must in ("propagate_vector", "propagate_matrix", "propagate_vec_coupling_S",
                 "propagate_vec_coupling_S_doubles", "propagate_vec_DIA", "diagonalizeLPD",
                 "propagate_double_2DIR_control", "calc_S1", "absorption", "do_1DFFT",
                 "calc_2DIR", "nise_absorption_batched", "nise_2dir_batched",
                 "nise_abs_plan_create", "nise_abs_plan_adjoint"):
        assert must in names, must


def test_library_exports_every_declared_symbol():
    L = C.CDLL(abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"declared in include/nise_abi.h but not exported: {missing}"


def test_version_string():
    L = abi.lib()
    assert b"gfx950" in L.nise_version()
