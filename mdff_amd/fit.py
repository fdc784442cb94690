"""Batched force-field fit over the C-ABI (mdff_fit_*), with a
torch.autograd.Function so that the loss can sit in a PyTorch optimisation
loop.  The forward and the gradient run in the HIP library (one persistent
launch for all snapshots); torch only carries the parameter vector.

Parameter kinds (include/mdff_gpu.h MDFF_FIT_*): A, RHO, C of a type pair,
ALPHA of a polarisable type, TTB / TTC of the Tang-Toennies damping of the
field at a dipole of type t1 from charges of type t2."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import capi

KINDS = {"A": 0, "RHO": 1, "C": 2, "ALPHA": 3, "TTB": 4, "TTC": 5}
MAXPARAM = 32

_P = C.POINTER
_dp = _P(C.c_double)
_ip = _P(C.c_int32)


def _lib() -> C.CDLL:
    return bind_fit(capi.gpu_lib())


def bind_fit(lib: C.CDLL) -> C.CDLL:
    """Declare the mdff_fit_* signatures on a loaded library."""
    if not getattr(lib, "_fit_bound", False):
        lib.mdff_fit_create.argtypes = [_P(capi.MdffDesc), C.c_int32, _ip, _dp, _dp, _ip, _ip, _dp, C.c_int32, _ip,
                                        _P(C.c_void_p)]
        lib.mdff_fit_set_targets.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
        lib.mdff_fit_set_launch.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        lib.mdff_fit_eval.argtypes = [C.c_void_p, _dp, C.c_int32, C.c_int32, _dp, _dp, _dp, _dp, _ip]
        lib.mdff_fit_destroy.argtypes = [C.c_void_p]
        lib.mdff_fit_last_error.argtypes = [C.c_void_p]
        lib.mdff_fit_last_error.restype = C.c_char_p
        lib._fit_bound = True
    return lib


def _a(x, dtype):
    return np.ascontiguousarray(x, dtype=dtype)


@dataclass
class FitResult:
    loss: float
    grad: np.ndarray | None
    forces: np.ndarray | None
    stress: np.ndarray | None
    nscf: np.ndarray | None


class FitProblem:
    """All snapshots of a fit in one device context.

    snapshots: mdff_amd.configs.System list sharing one force field (the
    descriptor of the first one provides types, charges, TT orders, SCF
    controls and the parameters that are not fitted); params: list of
    (kind, t1, t2) with kind in KINDS."""

    def __init__(self, snapshots, params, scf_tol: float | None = None, max_iter: int | None = None,
                 lib: C.CDLL | None = None):
        import copy
        self.lib = bind_fit(lib) if lib is not None else _lib()
        d = copy.deepcopy(snapshots[0].desc)
        if scf_tol is not None:
            d.conv_tol_ind = scf_tol
        if max_iter is not None:
            d.max_scf_pol_iter = max_iter
        self.desc = d
        self.params = [(k, int(t1), int(t2)) for (k, t1, t2) in params]
        self.nsnap = len(snapshots)
        self.natm = np.array([s.natm for s in snapshots], np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(self.natm)])
        cells = np.concatenate([np.asarray(s.desc.cell[:], float) for s in snapshots])
        cut = np.array([[s.desc.cutshortrange, s.desc.cutlongrange, s.desc.alphaES] for s in snapshots], float)
        kmax = np.array([list(s.desc.kES) for s in snapshots], np.int32)
        itype = np.concatenate([s.itype for s in snapshots]).astype(np.int32)
        pos = np.concatenate([s.pos for s in snapshots]).astype(float)
        spec = np.array([[KINDS[k], t1, t2] for (k, t1, t2) in self.params], np.int32).reshape(-1)
        self._keep = [_a(cells, float), _a(cut, float), _a(kmax, np.int32), _a(itype, np.int32), _a(pos, float),
                      _a(spec, np.int32), _a(self.natm, np.int32)]
        h = C.c_void_p()
        rc = self.lib.mdff_fit_create(C.byref(d), self.nsnap, self._keep[6].ctypes.data_as(_ip),
                                      self._keep[0].ctypes.data_as(_dp), self._keep[1].ctypes.data_as(_dp),
                                      self._keep[2].ctypes.data_as(_ip), self._keep[3].ctypes.data_as(_ip),
                                      self._keep[4].ctypes.data_as(_dp), len(self.params),
                                      self._keep[5].ctypes.data_as(_ip), C.byref(h))
        if rc != 0:
            msg = self.lib.mdff_fit_last_error(h).decode() if h.value else ""
            if h.value:
                self.lib.mdff_fit_destroy(h)
            raise capi.MdffError(rc, "mdff_fit_create", msg)
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.mdff_fit_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, where: str) -> None:
        if rc != 0:
            raise capi.MdffError(rc, where, self.lib.mdff_fit_last_error(self.h).decode())

    def set_targets(self, forces: np.ndarray, stress: np.ndarray, w_stress: float) -> None:
        f = _a(forces, float).reshape(-1)
        t = _a(stress, float).reshape(-1)
        if f.size != 3 * int(self.offsets[-1]) or t.size != 9 * self.nsnap:
            raise ValueError("target shapes")
        self._check(self.lib.mdff_fit_set_targets(self.h, f.ctypes.data_as(_dp), t.ctypes.data_as(_dp),
                                                  float(w_stress)), "set_targets")

    def set_launch(self, block: int, grid: int) -> None:
        self._check(self.lib.mdff_fit_set_launch(self.h, int(block), int(grid)), "set_launch")

    def eval(self, theta, precision: int = capi.PREC_MIXED, grad: bool = True, outputs: bool = False) -> FitResult:
        th = _a(theta, float).reshape(-1)
        if th.size != len(self.params):
            raise ValueError("theta size")
        loss = C.c_double()
        g = np.zeros(max(len(self.params), 1))
        F = np.zeros(3 * int(self.offsets[-1])) if outputs else None
        tau = np.zeros(9 * self.nsnap) if outputs else None
        ns = np.zeros(2 * self.nsnap, np.int32) if outputs else None
        rc = self.lib.mdff_fit_eval(self.h, th.ctypes.data_as(_dp), int(precision), 1 if grad else 0,
                                    C.byref(loss), g.ctypes.data_as(_dp),
                                    F.ctypes.data_as(_dp) if outputs else None,
                                    tau.ctypes.data_as(_dp) if outputs else None,
                                    ns.ctypes.data_as(_ip) if outputs else None)
        self._check(rc, "mdff_fit_eval")
        return FitResult(loss.value, g[: len(self.params)].copy() if grad else None,
                         F.reshape(-1, 3) if outputs else None, tau.reshape(-1, 9) if outputs else None,
                         ns.reshape(-1, 2) if outputs else None)

    def autotune(self, theta, precision: int = capi.PREC_MIXED, blocks=(64, 128, 256), reps: int = 2) -> dict:
        """Pick the block size (atoms and k-vectors per workgroup of the phase
        launches) with the lowest measured time of a full forward + gradient
        evaluation; keeps the winner."""
        import time
        best, table = None, {}
        for b in blocks:
            self.set_launch(b, 1)
            self.eval(theta, precision)                   # warm
            t0 = time.perf_counter()
            for _ in range(reps):
                self.eval(theta, precision)
            dt = (time.perf_counter() - t0) / reps
            table[b] = dt
            if best is None or dt < best[0]:
                best = (dt, b)
        self.set_launch(best[1], 1)
        return {"block": best[1], "seconds": best[0], "table": {str(b): v for b, v in table.items()}}


def theta_of(params, table: dict) -> np.ndarray:
    """Parameter vector from a {(kind, t1, t2): value} table (ALPHA uses t2 = 0)."""
    return np.array([table[(k, t1, t2)] for (k, t1, t2) in params], float)


def make_fit_loss():
    """torch.autograd.Function whose forward is the fit loss of a FitProblem
    and whose backward is the library's adjoint gradient."""
    import torch

    class FitLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, problem, precision):
            th = theta.detach().to("cpu", torch.float64).numpy()
            r = problem.eval(th, precision, grad=theta.requires_grad)
            if r.grad is not None:
                ctx.save_for_backward(torch.as_tensor(r.grad, dtype=theta.dtype, device=theta.device))
            else:
                ctx.save_for_backward(torch.zeros_like(theta))
            return torch.as_tensor(r.loss, dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g, None, None

    return FitLoss


def plain_targets(snapshots, precision: int = capi.PREC_FP64, scf_tol: float | None = None,
                  max_iter: int | None = None) -> tuple[np.ndarray, np.ndarray]:
    """Forces (concatenated, N x 3) and stress tau_nonb + tau_coul (nsnap x 9,
    GPa) of every snapshot from the single-configuration engforce (the plain
    method), at the parameters in the snapshots' descriptors."""
    import copy
    F, T = [], []
    for s in snapshots:
        d = copy.deepcopy(s.desc)
        d.precision = precision
        if scf_tol is not None:
            d.conv_tol_ind = scf_tol
        if max_iter is not None:
            d.max_scf_pol_iter = max_iter
        d.calc_md = 0
        eng = capi.gpu_engine(d, s.itype)
        try:
            eng.set_positions(s.pos)
            f, t = eng.engforce()
        finally:
            eng.close()
        F.append(f)
        T.append(np.ravel(t["tau_nonb"]) + np.ravel(t["tau_coul"]))
    return np.concatenate(F), np.array(T)


def with_theta(snapshots, params, theta) -> list:
    """Copies of the snapshots with the parameter vector written into their
    descriptors (for the plain method at theta)."""
    import copy
    out = []
    for s in snapshots:
        c = copy.copy(s)
        d = copy.deepcopy(s.desc)
        for (k, t1, t2), v in zip(params, theta):
            if k in ("A", "RHO", "C"):
                field = {"A": "Abmhftd", "RHO": "Bbmhftd", "C": "Cbmhftd"}[k]
                val = 1.0 / v if k == "RHO" else v
                getattr(d, field)[capi.t2(t1, t2)] = val
                getattr(d, field)[capi.t2(t2, t1)] = val
            elif k == "ALPHA":
                for a in range(3):
                    d.poldip[(t1 - 1) + 16 * a + 48 * a] = v
            else:
                field = "pol_damp_b" if k == "TTB" else "pol_damp_c"
                getattr(d, field)[capi.t3(t1, t1, t2)] = v
                getattr(d, field)[capi.t3(t1, t2, t1)] = v
        c.desc = d
        out.append(c)
    return out
