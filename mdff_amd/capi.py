"""ctypes view of the C-ABI in include/mdff_gpu.h.

The structures below are field-for-field copies of ``mdff_desc`` and
``mdff_terms``; ``Engine`` wraps one ``mdff_ctx``.  The same class drives any
shared library exporting the same entry points under another prefix, which is
how the tests run the CPU restatement side by side with the GPU library.

Reference boundary: SUBROUTINE engforce, src/engforce_driver.f90:102-190.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

NTYPEMAX = 16
ABI_VERSION = 3

STATUS = {
    0: "MDFF_OK",
    1: "MDFF_ERR_BAD_ARG",
    2: "MDFF_ERR_LIST_OVERFLOW",
    3: "MDFF_ERR_SCF_NOT_CONVERGED",
    4: "MDFF_ERR_POLAR_CATASTROPHE",
    5: "MDFF_ERR_HIP",
    6: "MDFF_ERR_NO_DEVICE",
    7: "MDFF_ERR_UNSUPPORTED",
}

PREC_FP64 = 0
PREC_MIXED = 1

# reference constants (src/constants.f90:42-50)
COUL_UNIT = 14.3996441494161
PRESS_UNIT = 0.00624150964712042
BOLTZ_UNIT = 8.61734229648141e-05
TIME_UNIT = 98.2269514139276


class MdffDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("natm", C.c_int32), ("ntype", C.c_int32),
        ("precision", C.c_int32), ("cell", C.c_double * 9),
        ("lnmlj", C.c_int32), ("lbmhft", C.c_int32), ("lbmhftd", C.c_int32),
        ("lmorse", C.c_int32), ("lcoulomb", C.c_int32), ("lsurf", C.c_int32),
        ("lvnlist", C.c_int32), ("calc_md", C.c_int32),
        ("cutshortrange", C.c_double), ("cutlongrange", C.c_double), ("skindiff", C.c_double),
        ("coul_unit", C.c_double), ("press_unit", C.c_double),
        ("trunc", C.c_int32), ("pad0", C.c_int32),
        ("epslj", C.c_double * 256), ("sigmalj", C.c_double * 256),
        ("plj", C.c_double * 256), ("qlj", C.c_double * 256),
        ("epsmor", C.c_double * 256), ("sigmamor", C.c_double * 256), ("rhomor", C.c_double * 256),
        ("Abmhftd", C.c_double * 256), ("Bbmhftd", C.c_double * 256), ("Cbmhftd", C.c_double * 256),
        ("Dbmhftd", C.c_double * 256), ("BDbmhftd", C.c_double * 256),
        ("alphaES", C.c_double), ("kES", C.c_int32 * 3), ("doefield", C.c_int32),
        ("mass", C.c_double * 16), ("qch", C.c_double * 16), ("dip", C.c_double * 48),
        ("ldip_polar", C.c_int32 * 16), ("poldip", C.c_double * 144),
        ("ldip_damping", C.c_int32 * 4096), ("pol_damp_b", C.c_double * 4096),
        ("pol_damp_c", C.c_double * 4096), ("pol_damp_k", C.c_int32 * 4096),
        ("conv_tol_ind", C.c_double), ("min_scf_pol_iter", C.c_int32),
        ("max_scf_pol_iter", C.c_int32), ("extrapolate_order", C.c_int32),
        ("algo_moment_from_pola", C.c_int32), ("omegakO", C.c_double),
        ("doefg", C.c_int32), ("pad1", C.c_int32),
    ]


class MdffTerms(C.Structure):
    _fields_ = [
        ("u_lj", C.c_double), ("u_bmhft", C.c_double), ("u_morse", C.c_double),
        ("vir_lj", C.c_double), ("vir_bmhft", C.c_double), ("vir_morse", C.c_double),
        ("u_coul", C.c_double), ("u_pol", C.c_double),
        ("u_dir", C.c_double), ("u_rec", C.c_double), ("u_self", C.c_double), ("u_surf", C.c_double),
        ("pvirial_coul", C.c_double), ("u_tail", C.c_double), ("p_tail", C.c_double),
        ("tau", C.c_double * 9), ("tau_nonb", C.c_double * 9), ("tau_coul", C.c_double * 9),
        ("n_scf", C.c_int32), ("list_rebuilt", C.c_int32), ("rmsd", C.c_double),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = np.array(v[:]).reshape(3, 3, order="F") if name.startswith("tau") else v
        return out


def t2(it: int, jt: int) -> int:
    """Index of type pair (it, jt), 1-based, in a 16x16 column-major table."""
    return (it - 1) + 16 * (jt - 1)


def t3(k: int, i: int, j: int) -> int:
    """Index of (k, i, j), 1-based, in a 16x16x16 column-major table."""
    return (k - 1) + 16 * (i - 1) + 256 * (j - 1)


def default_desc(natm: int, ntype: int, cell: np.ndarray) -> MdffDesc:
    """Descriptor with the reference defaults (src/control.f90:201-245,
    src/non_bonded.f90:206-240, src/coulomb.f90:160-180, src/pim.f90:107-140)."""
    d = MdffDesc()
    d.abi_version = ABI_VERSION
    d.natm = natm
    d.ntype = ntype
    d.precision = PREC_FP64
    # cell rows are the lattice vectors (POSFF order); column-major A(:,j) = vector j,
    # so the row-major flattening of the rows is the Fortran storage of A
    d.cell[:] = list(np.asarray(cell, dtype=np.float64).reshape(9))
    d.lvnlist = 1
    d.calc_md = 1
    d.skindiff = 0.15
    d.coul_unit = COUL_UNIT
    d.press_unit = PRESS_UNIT
    for i in range(256):
        d.plj[i] = 6.0
        d.qlj[i] = 12.0
    d.alphaES = 1.0
    d.kES[:] = [10, 10, 10]
    d.conv_tol_ind = 1e-6
    d.min_scf_pol_iter = 3
    d.max_scf_pol_iter = 100
    d.extrapolate_order = 0
    d.algo_moment_from_pola = 0
    d.omegakO = 0.7
    return d


def set_pair(d: MdffDesc, field: str, it: int, jt: int, value: float) -> None:
    """Set a symmetric type-pair parameter (symmetric_pot, src/non_bonded.f90:281-296)."""
    arr = getattr(d, field)
    arr[t2(it, jt)] = value
    arr[t2(jt, it)] = value


class MdffError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str = ""):
        self.code = code
        super().__init__(f"{where}: {STATUS.get(code, code)} {msg}".strip())


_P = C.POINTER


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_P(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(_P(C.c_int32))


def _bind(lib: C.CDLL, prefix: str) -> None:
    dp, ip, vp = _P(C.c_double), _P(C.c_int32), C.c_void_p
    sig = {
        "ctx_create" if prefix == "mdff_" else "create": ([_P(MdffDesc), ip, _P(vp)], C.c_int),
        "set_positions": ([vp, dp, dp, dp], C.c_int),
        "get_positions": ([vp, dp, dp, dp], C.c_int),
        "set_velocities": ([vp, dp, dp, dp], C.c_int),
        "get_velocities": ([vp, dp, dp, dp], C.c_int),
        "get_forces": ([vp, dp, dp, dp], C.c_int),
        "set_itime": ([vp, C.c_int32], C.c_int),
        "vnlist_build": ([vp], C.c_int),
        "vnlistcheck": ([vp, ip], C.c_int),
        "get_pairs": ([vp, C.c_int32, ip, ip, C.c_int64, _P(C.c_int64)], C.c_int),
        "engforce": ([vp, dp, dp, dp, _P(MdffTerms)], C.c_int),
        "get_dipoles": ([vp, dp, dp], C.c_int),
        "get_efg": ([vp, dp], C.c_int),
        "last_terms": ([vp, _P(MdffTerms)], C.c_int),
        "md_vv": ([vp, C.c_int32, C.c_double, dp, dp], C.c_int),
        "md_nhcn": ([vp, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32,
                     C.c_int32, dp, dp, dp, dp, dp], C.c_int),
        "csvr_rescale": ([vp, C.c_double, C.c_double, C.c_double, C.c_double, dp], C.c_int),
        "calc_temp": ([vp, C.c_double, dp, dp], C.c_int),
        "rescale_velocities": ([vp, C.c_double], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, prefix + name)
        fn.argtypes = args
        fn.restype = res
    err = getattr(lib, prefix + "last_error")
    err.argtypes = [vp]
    err.restype = C.c_char_p


class Engine:
    """One context of a library exporting the mdff_* (or a same-shaped) API.

    Mirrors how the reference drives engforce: positions in, engforce,
    forces/energies/stress out (src/engforce_driver.f90:102-190)."""

    def __init__(self, lib: C.CDLL, prefix: str, desc: MdffDesc, itype: np.ndarray):
        self.lib, self.prefix = lib, prefix
        self.desc = desc
        self.n = int(desc.natm)
        self._itype = np.ascontiguousarray(itype, dtype=np.int32)
        if self._itype.shape != (self.n,):
            raise ValueError("itype must have natm entries")
        h = C.c_void_p()
        create = getattr(lib, prefix + ("ctx_create" if prefix == "mdff_" else "create"))
        rc = create(C.byref(desc), _ip(self._itype), C.byref(h))
        if rc != 0:
            raise MdffError(rc, "create")
        self.h = h
        self._destroy = getattr(lib, prefix + ("ctx_destroy" if prefix == "mdff_" else "destroy"))

    def close(self) -> None:
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name: str, *args) -> None:
        rc = getattr(self.lib, self.prefix + name)(self.h, *args)
        if rc != 0:
            msg = getattr(self.lib, self.prefix + "last_error")(self.h)
            raise MdffError(rc, name, msg.decode() if msg else "")

    def _vec3(self, name: str, x, y, z) -> None:
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z)]
        for v in a:
            if v.shape != (self.n,):
                raise ValueError(f"{name}: arrays must have natm entries")
        self._call(name, *[_dp(v) for v in a])

    def _get3(self, name: str) -> np.ndarray:
        out = np.zeros((3, self.n))
        self._call(name, _dp(out[0]), _dp(out[1]), _dp(out[2]))
        return out

    def set_positions(self, r: np.ndarray) -> None:
        r = np.asarray(r, dtype=np.float64)
        self._vec3("set_positions", r[:, 0], r[:, 1], r[:, 2])

    def get_positions(self) -> np.ndarray:
        return self._get3("get_positions").T.copy()

    def set_velocities(self, v: np.ndarray) -> None:
        v = np.asarray(v, dtype=np.float64)
        self._vec3("set_velocities", v[:, 0], v[:, 1], v[:, 2])

    def get_velocities(self) -> np.ndarray:
        return self._get3("get_velocities").T.copy()

    def get_forces(self) -> np.ndarray:
        return self._get3("get_forces").T.copy()

    def set_itime(self, itime: int) -> None:
        self._call("set_itime", C.c_int32(itime))

    def vnlist_build(self) -> None:
        self._call("vnlist_build")

    def vnlistcheck(self) -> bool:
        rb = C.c_int32(0)
        self._call("vnlistcheck", C.byref(rb))
        return bool(rb.value)

    def get_pairs(self, which: int) -> tuple[np.ndarray, np.ndarray]:
        n = C.c_int64(0)
        point = np.zeros(self.n + 1, dtype=np.int32)
        rc = getattr(self.lib, self.prefix + "get_pairs")(self.h, which, _ip(point),
                                                          _ip(np.zeros(1, np.int32)), 0, C.byref(n))
        if rc not in (0, 1):
            raise MdffError(rc, "get_pairs")
        lst = np.zeros(max(int(n.value), 1), dtype=np.int32)
        self._call("get_pairs", C.c_int32(which), _ip(point), _ip(lst), C.c_int64(lst.size), C.byref(n))
        return point, lst[: int(n.value)]

    def engforce(self) -> tuple[np.ndarray, dict]:
        t = MdffTerms()
        f = np.zeros((3, self.n))
        self._call("engforce", _dp(f[0]), _dp(f[1]), _dp(f[2]), C.byref(t))
        return f.T.copy(), t.as_dict()

    def last_terms(self) -> dict:
        t = MdffTerms()
        self._call("last_terms", C.byref(t))
        return t.as_dict()

    def get_dipoles(self) -> tuple[np.ndarray, np.ndarray]:
        mu = np.zeros((self.n, 3))
        ef = np.zeros((self.n, 3))
        self._call("get_dipoles", _dp(mu), _dp(ef))
        return mu, ef

    def get_efg(self) -> np.ndarray:
        """efg_t of the last force evaluation as (natm, 3, 3): [i, a, b] is
        component (a+1, b+1) of the reference array efg(3,3,natm)."""
        g = np.zeros((self.n, 3, 3))
        self._call("get_efg", _dp(g))
        # the C layout efg[9i + a + 3b] is column-major per atom
        return g.transpose(0, 2, 1).copy()

    def md_vv(self, nsteps: int, dt: float) -> tuple[np.ndarray, np.ndarray]:
        ek = np.zeros(nsteps)
        up = np.zeros(nsteps)
        self._call("md_vv", C.c_int32(nsteps), C.c_double(dt), _dp(ek), _dp(up))
        return ek, up

    def md_nhcn(self, nsteps: int, dt: float, temp: float, timesca: float, xi: np.ndarray,
                vxi: np.ndarray, yosh_order: int = 1, nhc_mults: int = 1):
        ek, up, env = np.zeros(nsteps), np.zeros(nsteps), np.zeros(nsteps)
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        vxi = np.ascontiguousarray(vxi, dtype=np.float64)
        self._call("md_nhcn", C.c_int32(nsteps), C.c_double(dt), C.c_double(temp),
                   C.c_double(timesca), C.c_int32(xi.size), C.c_int32(yosh_order),
                   C.c_int32(nhc_mults), _dp(xi), _dp(vxi), _dp(ek), _dp(up), _dp(env))
        return ek, up, env, xi, vxi

    def csvr_rescale(self, temp: float, taut: float, r1: float, sumnoise: float, conint: float) -> float:
        ci = C.c_double(conint)
        self._call("csvr_rescale", C.c_double(temp), C.c_double(taut), C.c_double(r1),
                   C.c_double(sumnoise), C.byref(ci))
        return ci.value

    def rescale_velocities(self, lam: float) -> None:
        self._call("rescale_velocities", C.c_double(lam))

    def calc_temp(self, boltz_unit: float = BOLTZ_UNIT) -> tuple[float, float]:
        t, e = C.c_double(), C.c_double()
        self._call("calc_temp", C.c_double(boltz_unit), C.byref(t), C.byref(e))
        return t.value, e.value


_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "lib" / "libmdff_gpu.so"
_gpu_lib = None


def gpu_lib() -> C.CDLL:
    """Load the in-tree HIP library.  There is no fallback: a missing library
    is an error (the GPU path is the product)."""
    global _gpu_lib
    if _gpu_lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        lib = C.CDLL(str(LIB_PATH))
        _bind(lib, "mdff_")
        for name in ("ctx_destroy",):
            getattr(lib, "mdff_" + name).argtypes = [C.c_void_p]
        lib.mdff_version.restype = C.c_char_p
        lib.mdff_last_timings.argtypes = [C.c_void_p, _P(C.c_double), _P(C.c_int32)]
        lib.mdff_set_cell.argtypes = [C.c_void_p, _P(C.c_double)]
        _gpu_lib = lib
    return _gpu_lib


def gpu_engine(desc: MdffDesc, itype: np.ndarray) -> Engine:
    return Engine(gpu_lib(), "mdff_", desc, itype)


def bind_same_api(lib: C.CDLL, prefix: str) -> C.CDLL:
    """Declare the signatures of a library exporting the same API under prefix."""
    _bind(lib, prefix)
    getattr(lib, prefix + "destroy").argtypes = [C.c_void_p]
    return lib


def last_timings(eng: Engine) -> np.ndarray:
    ms = (C.c_double * 8)()
    n = C.c_int32(0)
    rc = eng.lib.mdff_last_timings(eng.h, ms, C.byref(n))
    if rc != 0:
        raise MdffError(rc, "last_timings")
    return np.array(ms[: n.value])


def build_dir() -> Path:
    return _HERE / "lib"


def env_flag(name: str) -> bool:
    return os.environ.get(name, "") not in ("", "0")
