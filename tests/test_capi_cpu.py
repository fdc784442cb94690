"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, and refuses bad arguments before touching a device."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from mdff_amd import capi, configs

ROOT = Path(__file__).resolve().parent.parent


def header_symbols():
    text = (ROOT / "include" / "mdff_gpu.h").read_text()
    return sorted(set(re.findall(r"\b(mdff_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = capi.gpu_lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s


def test_desc_layout_matches_header_defaults():
    lib = capi.gpu_lib()
    lib.mdff_desc_defaults.argtypes = [C.POINTER(capi.MdffDesc)]
    d = capi.MdffDesc()
    lib.mdff_desc_defaults(C.byref(d))
    ref = capi.default_desc(10, 1, np.eye(3))
    assert d.abi_version == capi.ABI_VERSION
    assert d.skindiff == 0.15 and d.coul_unit == capi.COUL_UNIT and d.press_unit == capi.PRESS_UNIT
    assert d.plj[37] == 6.0 and d.qlj[255] == 12.0
    assert list(d.kES) == [10, 10, 10] and d.conv_tol_ind == 1e-6
    assert d.min_scf_pol_iter == 3 and d.max_scf_pol_iter == 100
    assert C.sizeof(capi.MdffDesc) == C.sizeof(type(ref))


def test_bad_arguments_are_rejected():
    lib = capi.gpu_lib()
    s = configs.two_charges()
    h = C.c_void_p()
    it = np.array([1, 3], np.int32)  # type out of range
    rc = lib.mdff_ctx_create(C.byref(s.desc), it.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(h))
    assert rc == 1
    d = s.desc
    d.lmorse = 1
    d.lbmhft = 1
    rc = lib.mdff_ctx_create(C.byref(d), s.itype.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(h))
    assert rc == 7
    d.lmorse = 0
    d.algo_moment_from_pola = 3
    rc = lib.mdff_ctx_create(C.byref(d), s.itype.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(h))
    assert rc == 7
