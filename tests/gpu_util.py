"""Helpers of the GPU parity tests: run the same system through the HIP
library and through the oracle, and measure the differences."""
from __future__ import annotations

import copy
import ctypes as C
import os
import sys

import numpy as np

import oracle_lib as O
from mdff_amd import capi

TERMS = ["u_lj", "u_bmhft", "u_morse", "u_coul", "u_pol", "u_dir", "u_rec", "u_self", "u_surf",
         "vir_lj", "vir_bmhft", "vir_morse", "u_tail", "p_tail"]


def note(msg: str) -> None:
    """Name the case before a GPU step (flushed, so a crash still shows it)."""
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def hip_runtime():
    """The HIP runtime the library links, as loaded with it (no second
    runtime in the process): device buffers for the entry points that take
    device pointers."""
    capi.gpu_lib()
    try:
        hip = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | os.RTLD_GLOBAL)
    except OSError:
        hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    for fn in (hip.hipMalloc, hip.hipMemcpy, hip.hipFree, hip.hipDeviceSynchronize):
        fn.restype = C.c_int
    return hip


def pair(system, precision=capi.PREC_FP64, **over):
    d = copy.deepcopy(system.desc)
    for k, v in over.items():
        setattr(d, k, v)
    d.precision = precision
    g = capi.gpu_engine(d, system.itype)
    d2 = copy.deepcopy(system.desc)
    for k, v in over.items():
        setattr(d2, k, v)
    o = O.oracle(d2, system.itype)
    for e in (g, o):
        e.set_positions(system.pos)
        e.set_velocities(system.vel)
    return g, o


def rel_force_err(f, fref):
    """max |dF| / max |F| over components (the FP64 parity metric)."""
    return float(np.abs(f - fref).max() / max(np.abs(fref).max(), 1e-300))


def rms_force_err(f, fref):
    """max_i |F_i - F_i^ref| / RMS_i |F_i^ref| (the mixed form's force-error
    metric: the worst atom's error against the typical force)."""
    dF = np.linalg.norm(np.asarray(f) - np.asarray(fref), axis=1)
    rms = float(np.sqrt(np.mean(np.sum(np.asarray(fref) ** 2, axis=1))))
    return float(dF.max() / max(rms, 1e-300))


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def compare_terms(tg, to, tol, keys=TERMS):
    bad = {}
    for k in keys:
        a, b = tg[k], to[k]
        if abs(b) < 1e-300 and abs(a) < 1e-300:
            continue
        e = abs(a - b) / max(abs(b), 1e-12)
        if e > tol:
            bad[k] = (a, b, e)
    for k in ("tau_nonb", "tau_coul", "tau"):
        if np.abs(to[k]).max() > 0 and rel(tg[k], to[k]) > tol:
            bad[k] = (tg[k], to[k], rel(tg[k], to[k]))
    return bad
