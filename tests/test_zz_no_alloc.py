"""mdff_calc_temp_noalloc: the kinetic energy and temperature of velocities
already in device memory, with no allocation (a fixed device array of the
library holds its block partials), against mdff_calc_temp on the same
velocities: two calls in a row, the second on fewer atoms."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from gpu_util import hip_runtime, note
from mdff_amd import capi, configs

pytestmark = pytest.mark.gpu


def _bind(lib):
    f = lib.mdff_calc_temp_noalloc
    f.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


def _to_device(hip, arr):
    arr = np.ascontiguousarray(arr, dtype=np.float64)
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), arr.nbytes) == 0
    assert hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0      # host to device
    return p


def test_calc_temp_noalloc_matches_calc_temp():
    """Device buffers come from the HIP runtime the library itself links, so
    the test needs no other GPU framework in the process."""
    f = _bind(capi.gpu_lib())
    hip = hip_runtime()
    for natm in (24000, 900):
        s = configs.sio2(natm=natm, seed=40 + natm % 7)
        note(f"calc_temp_noalloc n={natm}")
        e = capi.gpu_engine(s.desc, s.itype)
        e.set_velocities(s.vel)
        t_ref, k_ref = e.calc_temp()
        e.close()
        masses = np.array([s.desc.mass[t - 1] for t in s.itype], dtype=np.float64)
        bufs = [_to_device(hip, x) for x in (masses, s.vel[:, 0], s.vel[:, 1], s.vel[:, 2], np.full(2, -1.0))]
        try:
            rc = f(natm, *bufs[:4], capi.BOLTZ_UNIT, None, bufs[4])
            assert hip.hipDeviceSynchronize() == 0
            assert rc == 0
            out = np.zeros(2)
            assert hip.hipMemcpy(out.ctypes.data, bufs[4], out.nbytes, 2) == 0   # device to host
        finally:
            for b in bufs:
                hip.hipFree(b)
        k, t = out
        assert abs(k - k_ref) <= 1e-13 * abs(k_ref), (k, k_ref)
        assert abs(t - t_ref) <= 1e-13 * abs(t_ref), (t, t_ref)


def test_calc_temp_noalloc_rejects_bad_arguments():
    f = _bind(capi.gpu_lib())
    assert f(0, 1, 1, 1, 1, capi.BOLTZ_UNIT, None, 1) == 1
    assert f(10, None, 1, 1, 1, capi.BOLTZ_UNIT, None, 1) == 1
    assert f(10, 1, 1, 1, 1, 0.0, None, 1) == 1
