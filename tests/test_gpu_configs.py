"""The benchmark workloads at (or near) their full sizes against the oracle,
through the C-ABI.

- NaCl rock salt, 512 ions: 100 velocity-Verlet steps each followed by the
  CSVR rescale (src/kinetic.f90:122-198) with the same caller noise.
- a-SiO2, 24,000 ions: the FP64 form's forces, terms and pair sets.
- LiF-like molten salt: a few thousand ions against the oracle (kmax high
  enough that the mixed form's structure factor runs on the FP32 matrix
  cores), a few NHC-3 steps; and the 200,000-ion workload's invariants.
- MgO-like PIM oxide (4,608 ions): SCF iteration counts, dipoles, forces.
"""
from __future__ import annotations

import copy

import numpy as np
import pytest

from gpu_util import compare_terms, note, pair, rel, rel_force_err, rms_force_err
from mdff_amd import capi, configs

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL_MIXED = 1e-5


def _lists_equal(g, o, which):
    pg, lg = g.get_pairs(which)
    po, lo = o.get_pairs(which)
    assert np.array_equal(pg, po)
    assert np.array_equal(lg, lo)
    return len(lo)


def test_nacl_512_csvr_100_steps_match_oracle():
    """cfg1: 512 ions, BMH + Ewald (alpha 0.3, rc 8, kmax 8), 100 steps of
    velocity Verlet + list check + CSVR rescale (taut 0: the kinetic energy
    resampled every step) with seeded caller noise; energies every step and
    the final state against the oracle."""
    s = configs.nacl(ncell=4)
    note(f"nacl csvr 100 steps n={s.natm}")
    g, o = pair(s)
    g.engforce()
    o.engforce()
    rng = np.random.default_rng(11)
    ndeg = 3 * s.natm - 3
    ci_g = ci_o = 0.0
    for step in range(100):
        eg, ug = g.md_vv(1, s.dt)
        eo, uo = o.md_vv(1, s.dt)
        assert abs(eg[0] - eo[0]) < 1e-9 * abs(eo[0]), step
        assert abs(ug[0] - uo[0]) < 1e-9 * abs(uo[0]), step
        r1 = float(rng.standard_normal())
        sn = float(rng.chisquare(ndeg - 1))
        ci_g = g.csvr_rescale(s.temp, 0.0, r1, sn, ci_g)
        ci_o = o.csvr_rescale(s.temp, 0.0, r1, sn, ci_o)
        assert abs(ci_g - ci_o) < 1e-9 * max(1.0, abs(ci_o)), step
    assert rel(g.get_positions(), o.get_positions()) < 1e-10
    assert rel(g.get_velocities(), o.get_velocities()) < 1e-8
    tg, to = g.calc_temp(), o.calc_temp()
    assert abs(tg[0] - to[0]) < 1e-8 * to[0]
    for which in (0, 1):
        _lists_equal(g, o, which)


def test_sio2_24k_fp64_matches_oracle():
    """cfg2 at full size: the FP64 form's forces (1e-10 of max |F|), energy
    and stress terms (1e-10) and both pair sets (bit-exact) against the
    oracle."""
    s = configs.sio2()
    note(f"sio2 24k fp64 vs oracle n={s.natm}")
    g, o = pair(s)
    fg, tg = g.engforce()
    fo, to = o.engforce()
    assert rel_force_err(fg, fo) < TOL64, rel_force_err(fg, fo)
    bad = compare_terms(tg, to, TOL64)
    assert not bad, bad
    for which in (0, 1):
        assert _lists_equal(g, o, which) > 0


def _lif_small():
    # 4,000 ions (L = 36.3 A); kmax 32 puts Kx + Ky + 2 = 66 staged X / Y
    # columns beyond the split-half structure factor's 64, so the mixed form
    # runs the FP32 matrix-core one (as the 200,000-ion workload at kmax 56)
    return configs.lif(natm=4000, seed=33, kmax=32)


def test_lif_4000_fp64_and_mixed_match_oracle():
    """cfg3 shape at 4,000 ions: FP64 within 1e-10, mixed within the 1e-5 bar
    (max_i |dF_i| / RMS |F|) of the oracle; Li-Li pairs (A = 0) skipped as
    the reference does (src/non_bonded.f90:1153)."""
    s = _lif_small()
    note(f"lif fp64/mixed vs oracle n={s.natm} kmax={s.desc.kES[0]}")
    g, o = pair(s)
    fo, to = o.engforce()
    fg, tg = g.engforce()
    assert rel_force_err(fg, fo) < TOL64
    bad = compare_terms(tg, to, TOL64)
    assert not bad, bad
    gm, _ = pair(s, precision=capi.PREC_MIXED)
    fm, tm = gm.engforce()
    assert rms_force_err(fm, fo) < TOL_MIXED, rms_force_err(fm, fo)
    assert abs(tm["u_coul"] - to["u_coul"]) < 1e-6 * abs(to["u_coul"])


def test_lif_4000_nhc3_steps_match_oracle():
    """Nose-Hoover chain of length 3 (src/integration.f90:538-701), 10 steps
    at 1200 K: per-step energies, thermostat variables and the final state
    against the oracle."""
    s = _lif_small()
    note(f"lif nhc-3 vs oracle n={s.natm}")
    g, o = pair(s)
    g.engforce()
    o.engforce()
    tau = s.extra["timesca_thermo"]
    r = g.md_nhcn(10, s.dt, s.temp, tau, np.zeros(3), np.zeros(3))
    q = o.md_nhcn(10, s.dt, s.temp, tau, np.zeros(3), np.zeros(3))
    for a, b in zip(r, q):
        assert rel(a, b) < 1e-9
    assert rel(g.get_positions(), o.get_positions()) < 1e-10
    assert rel(g.get_velocities(), o.get_velocities()) < 1e-8


def test_lif_200k_invariants_and_nhc3():
    """cfg3 at full size (200,000 ions, kmax 56): the forces sum to zero,
    the mixed form is within the bar of the FP64 form, and NHC-3 steps of
    the mixed form keep the extended energy."""
    s = configs.lif()
    note(f"lif 200k invariants n={s.natm} kmax={s.desc.kES[0]}")
    g64 = capi.gpu_engine(copy.deepcopy(s.desc), s.itype)
    g64.set_positions(s.pos)
    f64, _ = g64.engforce()
    g64.close()
    fmax = np.abs(f64).max()
    assert np.abs(f64.sum(0)).max() < 1e-8 * fmax * np.sqrt(s.natm)
    d = copy.deepcopy(s.desc)
    d.precision = capi.PREC_MIXED
    g32 = capi.gpu_engine(d, s.itype)
    g32.set_positions(s.pos)
    f32, _ = g32.engforce()
    assert rms_force_err(f32, f64) < TOL_MIXED, rms_force_err(f32, f64)
    # from the raw packing the forces are large: relax first, then thermostat
    pos, _ = configs.relax_sd(g32, s.pos, iters=100)
    g32.set_positions(pos)
    g32.set_velocities(s.vel)
    g32.engforce()
    ek, up, env, xi, vxi = g32.md_nhcn(20, s.dt, s.temp, s.extra["timesca_thermo"], np.zeros(3), np.zeros(3))
    H = ek + up + env
    note(f"lif 200k nhc-3: H {H[0]:.6e} -> {H[-1]:.6e}, ekin {ek[0]:.6e}")
    assert np.all(np.isfinite(H))
    assert abs(H[-1] - H[0]) < 1e-2 * abs(ek[0])


def test_lif_200k_relaxed_mixed_force_error():
    """A relaxed 200,000-ion LiF-like melt at kmax 30 (small forces, so the
    relative error is demanding): the mixed form within the bar of the FP64
    form.  Before the split-half structure factor alternated the sign of its
    atom tiles, an offset of the matrix cores' FP32 accumulation that grows
    with the atom count left 3.3e-5 here, at the ions near the cell's
    corner."""
    s = configs.lif(kmax=30)
    note(f"lif 200k relaxed n={s.natm} kmax={s.desc.kES[0]}")
    g64 = capi.gpu_engine(copy.deepcopy(s.desc), s.itype)
    pos, _ = configs.relax_sd(g64, s.pos, iters=200)
    g64.set_positions(pos)
    f64, _ = g64.engforce()
    g64.close()
    d = copy.deepcopy(s.desc)
    d.precision = capi.PREC_MIXED
    g32 = capi.gpu_engine(d, s.itype)
    g32.set_positions(pos)
    f32, _ = g32.engforce()
    g32.close()
    assert rms_force_err(f32, f64) < TOL_MIXED, rms_force_err(f32, f64)


def test_mgo_pim_fp64_matches_oracle():
    """cfg4's force field and SCF (tolerance 1e-7, TT damping, ASPC off) on
    8 x 8 x 9 cells, 4,608 ions (the oracle's SCF at 12,000 ions takes
    minutes on the host): the same SCF iteration count, dipoles, forces and
    terms as the oracle."""
    s = configs.mgo_pim(cells=(8, 8, 9))
    note(f"mgo pim fp64 vs oracle n={s.natm}")
    g, o = pair(s, calc_md=0)
    fg, tg = g.engforce()
    fo, to = o.engforce()
    assert tg["n_scf"] == to["n_scf"]
    assert rel_force_err(fg, fo) < TOL64
    bad = compare_terms(tg, to, TOL64)
    assert not bad, bad
    mg, _ = g.get_dipoles()
    mo, _ = o.get_dipoles()
    assert rel(mg, mo) < TOL64


def test_mgo_pim_12k_mixed_meets_the_force_bar():
    """cfg4 at its full 12,000 ions: the mixed form (FP32 real space,
    FP32-class reciprocal kernels with the FP64 inner k-box in the force
    evaluation) against the FP64 form, both on the GPU: max_i |dF_i| / RMS
    |F| within 1e-5, the same SCF iteration count."""
    s = configs.mgo_pim()
    note(f"mgo pim 12k mixed vs fp64 n={s.natm}")
    out = {}
    for prec in (capi.PREC_FP64, capi.PREC_MIXED):
        d = copy.deepcopy(s.desc)
        d.precision = prec
        e = capi.gpu_engine(d, s.itype)
        e.set_positions(s.pos)
        out[prec] = e.engforce()
        e.close()
    (f64, t64), (fm, tm) = out[capi.PREC_FP64], out[capi.PREC_MIXED]
    assert tm["n_scf"] == t64["n_scf"]
    assert rms_force_err(fm, f64) < 1e-5, rms_force_err(fm, f64)
