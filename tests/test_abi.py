"""The C-ABI library loads and exports every symbol include/lbm_mirror_amd/lbm.h
declares (CPU only: no compute call needs a GPU here)."""
import ctypes as C
import os
import re

import pytest

import lbm_mirror_amd as lm
from lbm_mirror_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "lbm_mirror_amd", "lbm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z_0-9]+)\s*\(", src)))


def test_header_and_python_export_list_agree():
    assert _declared() == sorted(lm.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(lm.LIB_PATH)
    for name in _declared():
        assert hasattr(L, name), name


def test_library_loads_and_reports_version():
    L = lm.load_library()
    assert b"gfx950" in L.lbm_version()
    assert L.lbm_device_count() >= 0


def test_struct_layouts_match_header():
    # offsets the C side relies on (lbm.h): 4 ints, then doubles
    assert C.sizeof(abi.OpDesc) == 4 * 4 + 8 * 4 + 8 + 8 * 3 + 8 + 8 * 5
    assert C.sizeof(abi.Particle) == 8 * (8 * 3 + 4)
    assert abi.ParticleParams.sub_steps.offset == 4 * 8 + 4 * 4


def test_argument_errors_need_no_gpu():
    L = lm.load_library()
    h = C.c_void_p()
    # bad lattice is rejected before any device call
    st = L.lbm_ctx_create(C.byref(h), 7, abi.F64, (C.c_uint64 * 3)(4, 4, 1), 0)
    assert st == abi.ERR_INVALID_ARG
    st = L.lbm_ctx_create(C.byref(h), abi.D2Q9, abi.F64, (C.c_uint64 * 3)(0, 4, 1), 0)
    assert st == abi.ERR_INVALID_ARG
    assert L.lbm_last_error()
    m, i = C.c_double(), C.c_double()
    assert L.lbm_spheroid_mass(2, abi.F64, 3.0, 1.5, C.byref(m), C.byref(i)) == abi.OK


def test_spheroid_mass_matches_oracle(orc):
    L = lm.load_library()
    for D in (2, 3):
        for st in (abi.F64, abi.F32):
            m, i = C.c_double(), C.c_double()
            L.lbm_spheroid_mass(D, st, 7.3, 1.5, C.byref(m), C.byref(i))
            assert (m.value, i.value) == orc.spheroid_mass(D, st, 7.3, 1.5)


def test_product_path_has_no_oracle_dependency():
    """The shipped package never imports or loads the oracle."""
    pkg = os.path.join(ROOT, "lbm_mirror_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".hpp", ".cpp", ".h")) or fn == "Makefile":
                txt = open(os.path.join(dirpath, fn), errors="ignore").read()
                assert "oracle" not in txt.replace("the oracle", ""), fn
