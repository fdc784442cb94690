"""Batched fit (mdff_fit_*): forward parity with the single-configuration
engforce, gradient against central finite differences of the same loss and
against finite differences of the oracle's loss (the reference gradient),
FP32 vs FP64, and the torch.autograd binding."""
from __future__ import annotations

import copy

import numpy as np
import pytest

from mdff_amd import capi, configs
from mdff_amd import fit as F

PARAMS = [("A", 1, 2), ("RHO", 1, 2), ("A", 2, 2), ("RHO", 2, 2), ("C", 2, 2), ("ALPHA", 2, 0),
          ("TTB", 2, 1), ("TTC", 2, 1), ("TTB", 2, 2), ("TTC", 2, 2)]
THETA_STAR = F.theta_of(PARAMS, configs.FIT_THETA_STAR)
# a start away from the hidden parameters (fixed pattern, a few per cent)
THETA0 = THETA_STAR * (1.0 + 0.03 * np.array([1, -1, 0.5, 0.7, -0.8, 1, -0.6, 0.9, 0.4, -0.5]))


def small_set(n=4, seed=11, nmax=300):
    return configs.fit_snapshots(n, nmin=96, nmax=nmax, seed=seed)


def test_fit_create_rejects_bad_arguments():
    lib = F._lib()
    import ctypes as C
    s = small_set(1)[0]
    h = C.c_void_p()
    nat = np.array([s.natm], np.int32)
    cells = np.asarray(s.desc.cell[:], float)
    cut = np.array([s.desc.cutshortrange, s.desc.cutlongrange, s.desc.alphaES])
    km = np.array(list(s.desc.kES), np.int32)
    pos = np.ascontiguousarray(s.pos)
    it = np.ascontiguousarray(s.itype, np.int32)
    spec = np.array([0, 1, 2], np.int32)
    dp, ip = F._dp, F._ip
    rc = lib.mdff_fit_create(C.byref(s.desc), 0, nat.ctypes.data_as(ip), cells.ctypes.data_as(dp),
                             cut.ctypes.data_as(dp), km.ctypes.data_as(ip), it.ctypes.data_as(ip),
                             pos.ctypes.data_as(dp), 1, spec.ctypes.data_as(ip), C.byref(h))
    assert rc == 1  # BAD_ARG: no snapshots
    rc = lib.mdff_fit_create(C.byref(s.desc), 1, nat.ctypes.data_as(ip), cells.ctypes.data_as(dp),
                             cut.ctypes.data_as(dp), km.ctypes.data_as(ip), it.ctypes.data_as(ip),
                             pos.ctypes.data_as(dp), 40, spec.ctypes.data_as(ip), C.byref(h))
    assert rc == 1  # BAD_ARG: more parameters than MDFF_FIT_MAXPARAM


@pytest.mark.gpu
def test_fit_forward_matches_engforce():
    """The batched forward at theta* equals the single-configuration FP64
    engforce of every snapshot (both SCFs converged to 1e-12)."""
    from gpu_util import note, rel
    snaps = small_set(5, seed=12, nmax=600)
    note(f"fit forward parity, sizes {[s.natm for s in snaps]}")
    P = F.FitProblem(snaps, PARAMS, scf_tol=1e-12, max_iter=400)
    Fref, Tref = F.plain_targets(snaps, scf_tol=1e-12, max_iter=400)
    P.set_targets(Fref, Tref, 1.0)
    r = P.eval(THETA_STAR, capi.PREC_FP64, grad=False, outputs=True)
    assert rel(r.forces, Fref) < 1e-9, rel(r.forces, Fref)
    assert rel(r.stress, Tref) < 1e-9, rel(r.stress, Tref)
    assert r.loss < 1e-12 * np.sum(Fref ** 2)
    P.close()


@pytest.fixture(scope="module")
def fd_problem():
    snaps = small_set(4, seed=13, nmax=260)
    Fref, Tref = F.plain_targets(snaps, scf_tol=1e-12, max_iter=400)
    P = F.FitProblem(snaps, PARAMS, scf_tol=1e-13, max_iter=600)
    P.set_targets(Fref, Tref, 0.05)
    return snaps, P, Fref, Tref


@pytest.mark.gpu
def test_fit_gradient_matches_finite_differences(fd_problem):
    """Adjoint gradient (FP64) vs central differences of the same loss,
    relative step 1e-5: agreement to 1e-6 of the gradient norm."""
    from gpu_util import note
    snaps, P, _, _ = fd_problem
    note(f"fit gradient vs FD, sizes {[s.natm for s in snaps]}")
    r = P.eval(THETA0, capi.PREC_FP64, grad=True)
    g_fd = np.zeros_like(THETA0)
    for k in range(len(THETA0)):
        h = 1e-5 * abs(THETA0[k])
        tp, tm = THETA0.copy(), THETA0.copy()
        tp[k] += h
        tm[k] -= h
        g_fd[k] = (P.eval(tp, capi.PREC_FP64, grad=False).loss - P.eval(tm, capi.PREC_FP64, grad=False).loss) / (2 * h)
    err = np.abs(r.grad - g_fd)
    scale = np.linalg.norm(g_fd)
    print("grad", r.grad, "fd", g_fd, "rel", err / np.maximum(np.abs(g_fd), 1e-300))
    assert np.linalg.norm(r.grad - g_fd) <= 1e-6 * scale
    # every component on its own where it is not negligible
    big = np.abs(g_fd) > 1e-6 * scale
    assert np.all(err[big] <= 1e-5 * np.abs(g_fd[big]) + 1e-7 * scale)


@pytest.mark.gpu
def test_fit_gradient_matches_reference_loss_differences():
    """The reference gradient: central differences of the loss evaluated with
    the oracle's engforce (FP64 C restatement of the reference) on two small
    snapshots, against the adjoint gradient of the batched engine."""
    import oracle_lib as O
    from gpu_util import note
    snaps = small_set(2, seed=14, nmax=130)
    note(f"fit gradient vs oracle FD, sizes {[s.natm for s in snaps]}")
    Fref, Tref = F.plain_targets(snaps, scf_tol=1e-12, max_iter=400)
    w = 0.05
    P = F.FitProblem(snaps, PARAMS, scf_tol=1e-13, max_iter=600)
    P.set_targets(Fref, Tref, w)
    g = P.eval(THETA0, capi.PREC_FP64, grad=True).grad

    def oracle_loss(theta):
        L = 0.0
        off = 0
        for s in F.with_theta(snaps, PARAMS, theta):
            d = copy.deepcopy(s.desc)
            d.conv_tol_ind = 1e-13
            d.max_scf_pol_iter = 600
            d.calc_md = 0
            o = O.oracle(d, s.itype)
            o.set_positions(s.pos)
            f, t = o.engforce()
            tau = np.ravel(t["tau_nonb"]) + np.ravel(t["tau_coul"])
            k = snaps.index(next(x for x in snaps if x.name == s.name))
            L += np.sum((f - Fref[off:off + s.natm]) ** 2) + w * np.sum((tau - Tref[k]) ** 2)
            off += s.natm
            o.close()
        return L

    g_ref = np.zeros_like(THETA0)
    for k in range(len(THETA0)):
        h = 1e-5 * abs(THETA0[k])
        tp, tm = THETA0.copy(), THETA0.copy()
        tp[k] += h
        tm[k] -= h
        g_ref[k] = (oracle_loss(tp) - oracle_loss(tm)) / (2 * h)
    print("grad", g, "oracle fd", g_ref)
    assert np.linalg.norm(g - g_ref) <= 1e-5 * np.linalg.norm(g_ref)


@pytest.mark.gpu
def test_fit_block_sizes_agree(fd_problem):
    """Atom and k blocks of 64 and 128 give the loss, gradient and forces of
    the 256-wide blocks up to the order of the per-snapshot block sums (1e-11
    relative)."""
    from gpu_util import note, rel
    snaps, P, _, _ = fd_problem
    note("fit block sizes 64 / 128 / 256")
    runs = {}
    try:
        for b in (256, 128, 64):
            P.set_launch(b, 1)
            runs[b] = P.eval(THETA0, capi.PREC_FP64, grad=True, outputs=True)
    finally:
        P.set_launch(256, 1)
    r0 = runs[256]
    for b in (128, 64):
        r = runs[b]
        assert abs(r.loss - r0.loss) <= 1e-11 * abs(r0.loss), b
        assert np.linalg.norm(r.grad - r0.grad) <= 1e-11 * np.linalg.norm(r0.grad), b
        assert rel(r.forces, r0.forces) < 1e-11, b


@pytest.mark.gpu
def test_fit_fp32_forward_and_gradient(fd_problem):
    """FP32 pair/table arithmetic: forces within 1e-5 of FP64 (relative to the
    largest force), gradient within 1e-3."""
    from gpu_util import note, rel
    snaps, P, _, _ = fd_problem
    note("fit fp32 vs fp64")
    r64 = P.eval(THETA0, capi.PREC_FP64, grad=True, outputs=True)
    r32 = P.eval(THETA0, capi.PREC_MIXED, grad=True, outputs=True)
    assert rel(r32.forces, r64.forces) < 1e-5
    assert np.linalg.norm(r32.grad - r64.grad) <= 1e-3 * np.linalg.norm(r64.grad)


@pytest.mark.gpu
def test_fit_torch_autograd_step(fd_problem):
    """FitLoss behind torch.autograd: backward returns the library gradient;
    a few gradient steps on the log-parameters reduce the loss."""
    import torch
    from gpu_util import note
    snaps, P, _, _ = fd_problem
    note("fit torch autograd")
    FitLoss = F.make_fit_loss()
    th = torch.tensor(THETA0, dtype=torch.float64, requires_grad=True)
    loss = FitLoss.apply(th, P, capi.PREC_FP64)
    loss.backward()
    ref = P.eval(THETA0, capi.PREC_FP64, grad=True)
    assert np.allclose(th.grad.numpy(), ref.grad, rtol=1e-12, atol=0)
    # steepest descent in log-parameters with a backtracking step
    logt = torch.log(torch.tensor(THETA0, dtype=torch.float64)).requires_grad_(True)
    l0 = FitLoss.apply(torch.exp(logt), P, capi.PREC_FP64)
    l0.backward()
    gdir = logt.grad.clone()
    step = 1e-2 / float(gdir.abs().max())
    for _ in range(20):
        with torch.no_grad():
            trial = torch.exp(logt - step * gdir)
        l1 = P.eval(trial.numpy(), capi.PREC_FP64, grad=False).loss
        if l1 < float(l0.detach()):
            break
        step *= 0.5
    assert l1 < float(l0.detach())
