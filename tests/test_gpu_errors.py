"""Error statuses of the C-ABI, forced from the host: conditions the reference
answers with a message and STOP (src/tools.f90:337-340,
src/moment_from_pola.f90:176-216) come back as status codes, with the same
decision as the oracle.  Every kernel exits normally: the list overflow is
found by the counting pass before any row is written, the catastrophe by the
check after a converged SCF."""
from __future__ import annotations

import numpy as np
import pytest

from gpu_util import note, pair
from mdff_amd import capi, configs

pytestmark = pytest.mark.gpu


def _dense_lj(n=12000, L=40.0, rc=19.0, seed=51):
    """LJ particles at 0.19 per A^3 with a 19 A cutoff: about 5,400
    neighbours per ion, more than the reference's 2 x 2000 list entries per
    ion (vnlmax, src/config.f90:43)."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, L, (n, 3))
    d = capi.default_desc(n, 1, configs.cubic(L))
    d.lnmlj = 1
    d.cutshortrange = rc
    d.skindiff = 0.1
    d.mass[0] = 40.0
    capi.set_pair(d, "epslj", 1, 1, 0.01)
    capi.set_pair(d, "sigmalj", 1, 1, 3.4)
    return configs.System("dense_lj", d, np.ones(n, np.int32), pos, np.zeros((n, 3)))


def test_list_overflow_status():
    s = _dense_lj()
    note(f"list overflow n={s.natm}")
    g = capi.gpu_engine(s.desc, s.itype)
    g.set_positions(s.pos)
    with pytest.raises(capi.MdffError) as e:
        g.vnlist_build()
    assert e.value.code == 2                       # MDFF_ERR_LIST_OVERFLOW
    with pytest.raises(capi.MdffError) as e:
        g.engforce()
    assert e.value.code == 2
    g.close()


def _catastrophe(pol):
    """A charge +1 and a neutral polarisable ion 2 A along +x in a 20 A box:
    mu_x = pol * E_x with E_x close to 1/4 e/A^2 (no damping)."""
    d = capi.default_desc(2, 2, configs.cubic(20.0))
    d.lcoulomb = 1
    d.cutlongrange = 9.0
    d.alphaES = 0.5
    d.kES[:] = [10, 10, 10]
    d.qch[0] = 1.0
    d.qch[1] = 0.0
    d.mass[0] = d.mass[1] = 16.0
    d.ldip_polar[1] = 1
    for a in range(3):
        d.poldip[1 + 16 * a + 48 * a] = pol
    d.conv_tol_ind = 1e-10
    pos = np.array([[5.0, 5.0, 5.0], [7.0, 5.0, 5.0]])
    return configs.System("catastrophe", d, np.array([1, 2], np.int32), pos, np.zeros((2, 3)))


def test_polarisation_catastrophe_status():
    """pol 50 A^3: mu_x about 12.5 e A > 10 -> MDFF_ERR_POLAR_CATASTROPHE on
    both sides; pol 20 A^3 (mu_x about 5) passes on both."""
    for pol, bad in ((50.0, True), (20.0, False)):
        s = _catastrophe(pol)
        note(f"catastrophe pol={pol}")
        g, o = pair(s)
        if bad:
            for eng in (g, o):
                with pytest.raises(capi.MdffError) as e:
                    eng.engforce()
                assert e.value.code == 4           # MDFF_ERR_POLAR_CATASTROPHE
        else:
            fg, tg = g.engforce()
            fo, to = o.engforce()
            mg, _ = g.get_dipoles()
            mo, _ = o.get_dipoles()
            assert 4.0 < mg[1, 0] < 10.0
            assert np.abs(mg - mo).max() < 1e-10 * np.abs(mo).max()
        g.close()
        o.close()


def test_scf_stops_at_max_iterations():
    """The reference leaves the SCF loop at max_scf_pol_iter without an
    error (src/moment_from_pola.f90:365-439): with a tolerance no iterate
    meets, both sides stop after the same number of iterations and agree."""
    s = configs.mgo_pim(cells=(2, 2, 3), alpha=0.8, rcut=4.2, kmax=8, tol=1e-30)
    s.desc.max_scf_pol_iter = 7
    note("scf at max iterations")
    g, o = pair(s)
    fg, tg = g.engforce()
    fo, to = o.engforce()
    assert tg["n_scf"] == to["n_scf"] == 7
    assert np.abs(fg - fo).max() < 1e-10 * np.abs(fo).max()
