"""The batched-fit kernels (mdff_amd/csrc/fit.hip) built for the host
(tests/emu: a workgroup runs as std::threads with a barrier) and checked on
the CPU: forward against the oracle's engforce, adjoint gradient against
central finite differences, FP32 against FP64.  The GPU tests
(tests/test_fit.py) repeat these checks on the device."""
from __future__ import annotations

import copy
import ctypes as C
import shutil

import numpy as np
import pytest

import oracle_lib as O
from mdff_amd import capi, configs
from mdff_amd import fit as F

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")

PB = [("A", 1, 2), ("RHO", 1, 2), ("A", 2, 2), ("RHO", 2, 2), ("C", 2, 2)]
PP = PB + [("ALPHA", 2, 0), ("TTB", 2, 1), ("TTC", 2, 1), ("TTB", 2, 2), ("TTC", 2, 2)]
PERTURB = np.array([1, -1, 0.5, 0.7, -0.8, 1, -0.6, 0.9, 0.4, -0.5])


@pytest.fixture(scope="module")
def cpu_lib():
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parent / "emu"))
    import build_fit_cpu
    return C.CDLL(str(build_fit_cpu.build()))


def _snaps(pol: bool):
    out = []
    for s in configs.fit_snapshots(1, nmin=96, nmax=100, seed=14):
        c = copy.copy(s)
        d = copy.deepcopy(s.desc)
        if not pol:
            d.ldip_polar[1] = 0
        d.conv_tol_ind = 1e-13
        d.max_scf_pol_iter = 400
        d.calc_md = 0
        c.desc = d
        out.append(c)
    return out


def _oracle_targets(snaps):
    Fs, Ts = [], []
    for s in snaps:
        o = O.oracle(copy.deepcopy(s.desc), s.itype)
        o.set_positions(s.pos)
        f, t = o.engforce()
        Fs.append(f)
        Ts.append(np.ravel(t["tau_nonb"]) + np.ravel(t["tau_coul"]))
        o.close()
    return np.concatenate(Fs), np.array(Ts)


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("pol", [False, True])
def test_fit_forward_cpu_build_matches_oracle(cpu_lib, pol):
    snaps = _snaps(pol)
    params = PP if pol else PB
    theta = F.theta_of(params, configs.FIT_THETA_STAR)
    Fr, Tr = _oracle_targets(snaps)
    P = F.FitProblem(snaps, params, scf_tol=1e-13, max_iter=400, lib=cpu_lib)
    P.set_targets(Fr, Tr, 1.0)
    P.set_launch(32, 1)
    r = P.eval(theta, capi.PREC_FP64, grad=False, outputs=True)
    assert _rel(r.forces, Fr) < 1e-11
    assert _rel(r.stress, Tr) < 1e-11
    r32 = P.eval(theta, capi.PREC_MIXED, grad=False, outputs=True)
    assert _rel(r32.forces, Fr) < 1e-5
    P.close()


def test_fit_adjoint_gradient_cpu_build_matches_finite_differences(cpu_lib):
    """Gradient through the converged SCF (implicit adjoint) against central
    differences of the same FP64 loss (relative step 1e-5): 1e-6 of the
    norm.  A planted error in any term (sign of the adjoint solve, a missing
    reciprocal part of h, a wrong TT derivative) fails this."""
    snaps = _snaps(True)
    theta0 = F.theta_of(PP, configs.FIT_THETA_STAR) * (1.0 + 0.03 * PERTURB)
    Fr, Tr = _oracle_targets(snaps)
    P = F.FitProblem(snaps, PP, scf_tol=1e-13, max_iter=400, lib=cpu_lib)
    P.set_targets(Fr, Tr, 0.05)
    P.set_launch(32, 1)
    g = P.eval(theta0, capi.PREC_FP64, grad=True).grad
    # one parameter of every kind (Buckingham A, rho, C; alpha; TT b, c)
    check = [0, 1, 4, 5, 6, 9]
    gfd = g.copy()
    for k in check:
        h = 1e-5 * abs(theta0[k])
        tp, tm = theta0.copy(), theta0.copy()
        tp[k] += h
        tm[k] -= h
        gfd[k] = (P.eval(tp, capi.PREC_FP64, grad=False).loss - P.eval(tm, capi.PREC_FP64, grad=False).loss) / (2 * h)
    assert np.linalg.norm((g - gfd)[check]) <= 1e-6 * np.linalg.norm(gfd[check])
    for k in check:
        assert abs(g[k] - gfd[k]) <= 1e-5 * abs(gfd[k]) + 1e-9 * np.linalg.norm(gfd[check]), (k, g[k], gfd[k])
    g32 = P.eval(theta0, capi.PREC_MIXED, grad=True).grad
    assert np.linalg.norm(g32 - g) <= 1e-3 * np.linalg.norm(g)
    P.close()


def test_fit_block_maps_cpu_build_agree(cpu_lib):
    """Several snapshots of different sizes cut into atom and k blocks of 32
    and of 64: the same loss, forces, iteration counts and gradient (FP64;
    only the order of the per-snapshot block sums differs)."""
    snaps = configs.fit_snapshots(3, nmin=96, nmax=216, seed=21)
    for s in snaps:
        s.desc.conv_tol_ind = 1e-12
        s.desc.max_scf_pol_iter = 200
    theta0 = F.theta_of(PP, configs.FIT_THETA_STAR) * (1.0 + 0.03 * PERTURB)
    Fr, Tr = _oracle_targets(snaps)
    P = F.FitProblem(snaps, PP, scf_tol=1e-12, max_iter=200, lib=cpu_lib)
    P.set_targets(Fr, Tr, 0.05)
    out = []
    for b in (32, 64):
        P.set_launch(b, 1)
        out.append(P.eval(theta0, capi.PREC_FP64, grad=True, outputs=True))
    r1, r2 = out
    assert abs(r2.loss - r1.loss) <= 1e-12 * abs(r1.loss)
    assert _rel(r2.forces, r1.forces) < 1e-12
    assert np.array_equal(r2.nscf, r1.nscf)
    assert np.linalg.norm(r2.grad - r1.grad) <= 1e-10 * np.linalg.norm(r1.grad)
    P.close()
