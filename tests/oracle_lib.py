"""Loads the oracle (CPU restatement, oracle/liboracle.so) and the reference's
own scalar routines (oracle/_ref/libmdff_ref.so) for the tests.  Test
infrastructure only."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

from mdff_amd.capi import Engine, MdffDesc, bind_same_api

ROOT = Path(__file__).resolve().parent.parent
ORACLE = ROOT / "oracle" / "liboracle.so"
REF = ROOT / "oracle" / "_ref" / "libmdff_ref.so"

_lib = None
_ref = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE.exists():
            raise RuntimeError(f"{ORACLE} missing: run make -C oracle")
        L = C.CDLL(str(ORACLE))
        bind_same_api(L, "orc_")
        L.orc_errfc.argtypes = [C.c_double]
        L.orc_errfc.restype = C.c_double
        L.orc_nint.argtypes = [C.c_double]
        L.orc_nint.restype = C.c_double
        dp = C.POINTER(C.c_double)
        L.orc_tt_damp.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, dp, dp]
        L.orc_lattice.argtypes = [dp, dp, dp, dp]
        L.orc_kmesh_count.argtypes = [C.POINTER(C.c_int32)]
        L.orc_kmesh.argtypes = [dp, C.c_double, C.POINTER(C.c_int32), dp, dp, dp, dp, dp, dp]
        L.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_cutoffs.argtypes = [C.c_void_p, dp, dp, dp, C.POINTER(C.c_int32)]
        L.orc_multipole_es.argtypes = [C.c_void_p, dp, C.POINTER(C.c_int32), C.c_int32, dp, dp, dp,
                                       dp, dp, dp, dp, dp, dp]
        _lib = L
    return _lib


def ref() -> C.CDLL | None:
    """The reference's errfc / TT routines built from its sources, if present."""
    global _ref
    if _ref is None and REF.exists():
        R = C.CDLL(str(REF))
        R.ref_errfc.argtypes = [C.c_double]
        R.ref_errfc.restype = C.c_double
        dp = C.POINTER(C.c_double)
        R.ref_tt_damp.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, dp, dp]
        _ref = R
    return _ref


def oracle(desc: MdffDesc, itype) -> Engine:
    return Engine(lib(), "orc_", desc, itype)
