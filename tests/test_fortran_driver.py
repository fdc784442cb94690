"""The Fortran host side: the iso_c_binding module and the driver program
(mdff_amd/fortran) read the reference's input files and call the C-ABI.

CPU: the driver parses a namelist control file and a POSFF and stops with the
library's status when no device is present.  GPU: the driver's t = 0 energies,
forces and a 5-step NVE trajectory match the oracle run on the same inputs."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import fortran_inputs as FI
from mdff_amd import capi, configs

pytestmark = pytest.mark.skipif(not FI.DRIVER.exists(), reason="Fortran driver not built (flang absent)")


def test_driver_binary_is_current():
    """The driver is a build output (not tracked): it must be newer than every
    Fortran source and the library it links, or the tests below would run a
    stale program."""
    t = FI.DRIVER.stat().st_mtime
    srcs = sorted((FI.ROOT / "mdff_amd" / "fortran").glob("*.f90"))
    stale = [p.name for p in srcs if p.stat().st_mtime > t]
    assert not stale, f"{FI.DRIVER} is older than {stale}: run __graft_entry__.build()"


def _run(workdir, *args, timeout=120):
    env = dict(os.environ)
    return subprocess.run([str(FI.DRIVER), "control.F", *args], cwd=workdir, capture_output=True, text=True,
                          timeout=timeout, env=env)


def test_driver_reads_reference_inputs(tmp_path):
    """Namelists with the reference's group and variable names and a POSFF
    are accepted; without a GPU the context creation reports NO_DEVICE."""
    s = configs.nacl(ncell=1)
    FI.write_inputs(tmp_path, s, {1: "Na", 2: "Cl"}, npas=0)
    r = _run(tmp_path)
    if r.returncode == 0:       # a GPU is present: the run itself is the check
        assert "u_coul" in r.stdout
        return
    assert "ERROR in mdff_ctx_create (status 6)" in r.stdout, r.stdout + r.stderr


def test_driver_rejects_unknown_tag(tmp_path):
    """A variable that is not in the reference's group stops the run, as the
    reference's iostat check does (src/control.f90:172-179)."""
    s = configs.nacl(ncell=1)
    FI.write_inputs(tmp_path, s, {1: "Na", 2: "Cl"}, extra_control=[" not_a_tag = 1"])
    r = _run(tmp_path)
    assert r.returncode != 0
    assert "controltag wrong tag" in r.stdout


@pytest.mark.gpu
def test_driver_engforce_and_nve_match_oracle(tmp_path):
    import oracle_lib as O
    from gpu_util import note, rel
    s = configs.nacl(ncell=2, seed=3)
    nsteps = 5
    FI.write_inputs(tmp_path, s, {1: "Na", 2: "Cl"}, npas=nsteps)
    note("fortran driver nacl ncell=2 5 NVE steps")
    r = _run(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    terms = FI.parse_terms(r.stdout)
    fdrv = np.loadtxt(tmp_path / "FORCES")

    o = O.oracle(s.desc, s.itype)
    o.set_positions(s.pos)
    o.set_velocities(s.vel)
    fo, to = o.engforce()
    for k in ("u_bmhft", "u_coul", "u_dir", "u_rec", "u_self", "vir_bmhft", "pvirial_coul"):
        assert abs(terms[k] - to[k]) <= 1e-10 * abs(to[k]), (k, terms[k], to[k])
    # the driver prints the 9 entries of the C struct; the tensors are symmetric
    assert rel(terms["tau_nonb"], np.ravel(to["tau_nonb"])) < 1e-10
    assert rel(terms["tau_coul"], np.ravel(to["tau_coul"])) < 1e-10
    assert rel(fdrv, fo) < 1e-10

    thermo = FI.parse_thermo(r.stdout)
    assert list(thermo["step"]) == list(range(nsteps + 1))
    ek, up = [], []
    for _ in range(nsteps):
        e, u = o.md_vv(1, 0.001 * capi.TIME_UNIT)
        ek.append(e[0])
        up.append(u[0])
    assert np.allclose(thermo["Utot"][1:], up, rtol=1e-9, atol=0)
    assert np.allclose(thermo["Ekin"][1:], ek, rtol=1e-8, atol=0)
    osz = FI.parse_osziff(tmp_path / "OSZIFF")
    assert osz.shape == (nsteps + 1, 13)
    # the same thermo in both outputs (E16.8 in OSZIFF)
    assert np.allclose(osz[:, 3], thermo["Ekin"], rtol=1e-7)
    assert np.allclose(osz[:, 4], thermo["Utot"], rtol=1e-7)
    assert np.allclose(osz[:, 12], thermo["Htot"], rtol=1e-7)
    cont = FI.read_contff(tmp_path / "CONTFF", s.natm)
    assert rel(cont[:, 0:3], o.get_positions()) < 1e-10
    assert rel(cont[:, 3:6], o.get_velocities()) < 1e-8


def _fold(pos, A):
    """kardir, periodicbc (nint, half away from zero), dirkar
    (src/lattice.f90:183-295): positions folded into the cell as TRAJFF
    writes them."""
    B = np.linalg.inv(A)          # rows: reciprocal vectors / omega
    sfr = pos @ B.T
    sfr = sfr - np.where(sfr >= 0, np.floor(sfr + 0.5), np.ceil(sfr - 0.5))
    return sfr @ A.T


@pytest.mark.gpu
def test_driver_trajff_and_osziff(tmp_path):
    """TRAJFF frames at the steps itime > itraj_start with itime % itraj_period
    == 0 (src/runmd.f90:349-353), positions folded into the cell, rvf data;
    OSZIFF records every fprint steps plus t = 0; the last frame is the
    final configuration of CONTFF."""
    from gpu_util import note, rel
    s = configs.nacl(ncell=2, seed=4)
    FI.write_inputs(tmp_path, s, {1: "Na", 2: "Cl"}, npas=6,
                    extra_control=[" ltraj = .true.", " itraj_start = 1", " itraj_period = 2",
                                   " trajff_data = 'rvf'"],
                    extra_md=[" fprint = 3", " nprint = 2"])
    note("fortran driver trajff / osziff")
    r = _run(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    frames = FI.read_trajff(tmp_path / "TRAJFF", s.natm, 9)
    assert len(frames) == 3                       # steps 2, 4, 6
    cont = FI.read_contff(tmp_path / "CONTFF", s.natm)
    A = np.array(s.desc.cell[:]).reshape(3, 3).T  # columns = lattice vectors
    assert rel(frames[-1][:, 0:3], _fold(cont[:, 0:3], A)) < 1e-10
    assert rel(frames[-1][:, 3:9], cont[:, 3:9]) < 1e-10
    osz = FI.parse_osziff(tmp_path / "OSZIFF")
    assert list(osz[:, 0].astype(int)) == [0, 3, 6]
    thermo = FI.parse_thermo(r.stdout)
    assert list(thermo["step"].astype(int)) == [0, 2, 4, 6]


@pytest.mark.gpu
def test_driver_restart_full_vs_split(tmp_path):
    """ex14 (examples/ex14/run_ex14.sh:13-51): a run of 2 n + 1 steps against
    a run of n steps that writes RESTART (lwrite_restart) followed by a
    full_restart run, which takes every parameter and the state from
    RESTART and runs from step n + 1 to step 2 n + 1 (itime1 = itime0 + npas,
    src/restart.f90:196-197).  The final configurations agree, and so do
    the thermo records of the steps both runs make."""
    from gpu_util import note, rel
    s = configs.nacl(ncell=2, seed=8, temp_K=600.0)
    n = 6
    full = tmp_path / "full"
    split = tmp_path / "split"
    FI.write_inputs(full, s, {1: "Na", 2: "Cl"}, npas=2 * n + 1)
    FI.write_inputs(split, s, {1: "Na", 2: "Cl"}, npas=n, extra_control=[" lwrite_restart = .true."])
    note("fortran driver restart: full run")
    r = _run(full)
    assert r.returncode == 0, r.stdout + r.stderr
    note("fortran driver restart: split run 1")
    r1 = _run(split)
    assert r1.returncode == 0, r1.stdout + r1.stderr
    assert (split / "RESTART").exists()
    FI.write_inputs(split, s, {1: "Na", 2: "Cl"}, npas=0, extra_control=[" full_restart = .true."])
    (split / "POSFF").unlink()                    # everything comes from RESTART
    note("fortran driver restart: split run 2")
    r2 = _run(split)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    cf = FI.read_contff(full / "CONTFF", s.natm)
    cs = FI.read_contff(split / "CONTFF", s.natm)
    assert rel(cs[:, 0:3], cf[:, 0:3]) < 1e-9
    assert rel(cs[:, 3:6], cf[:, 3:6]) < 1e-7
    tf = FI.parse_thermo(r.stdout)
    t2 = FI.parse_thermo(r2.stdout)
    assert list(t2["step"].astype(int)) == list(range(n, 2 * n + 2))
    sel = [int(k) for k in t2["step"]]
    assert np.allclose(t2["Utot"], tf["Utot"][sel], rtol=1e-9)
    assert np.allclose(t2["Ekin"], tf["Ekin"][sel], rtol=1e-8)


@pytest.mark.gpu
def test_driver_nhcn_matches_oracle(tmp_path):
    """integrator 'nvt-nhcn' (src/integration.f90:538-701) with a chain of 3
    at 600 K, timesca_thermo 0.05 ps: the driver's steps against the
    oracle's chain from the same start (xi = vxi = 0): per-step potential and
    kinetic energies and the final positions and velocities."""
    import oracle_lib as O
    from gpu_util import note, rel
    s = configs.nacl(ncell=2, seed=5, temp_K=600.0)
    nsteps = 6
    FI.write_inputs(tmp_path, s, {1: "Na", 2: "Cl"}, npas=nsteps,
                    extra_md=[" integrator = 'nvt-nhcn'", " temp = 600.0", " timesca_thermo = 0.05",
                              " nhc_n = 3"])
    note("fortran driver nvt-nhcn")
    r = _run(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    o = O.oracle(s.desc, s.itype)
    o.set_positions(s.pos)
    o.set_velocities(s.vel)
    o.engforce()
    ek, up, env, xi, vxi = o.md_nhcn(nsteps, 0.001 * capi.TIME_UNIT, 600.0 * capi.BOLTZ_UNIT,
                                      0.05 * capi.TIME_UNIT, np.zeros(3), np.zeros(3))
    thermo = FI.parse_thermo(r.stdout)
    assert list(thermo["step"].astype(int)) == list(range(nsteps + 1))
    assert np.allclose(thermo["Utot"][1:], up, rtol=1e-9, atol=0)
    assert np.allclose(thermo["Ekin"][1:], ek, rtol=1e-8, atol=0)
    cont = FI.read_contff(tmp_path / "CONTFF", s.natm)
    assert rel(cont[:, 0:3], o.get_positions()) < 1e-10
    assert rel(cont[:, 3:6], o.get_velocities()) < 1e-8


@pytest.mark.gpu
def test_driver_csvr_equilibration_replayed_on_oracle(tmp_path):
    """nve-vv with lcsvr and nequil = npas (src/runmd.f90:320-331): the driver
    prints the deviates of every CSVR rescale; none at itime0 = 1 and
    itime1 = npas (src/md.f90:308-309); replaying them through the oracle's
    rescale (src/kinetic.f90:122-198) between its VV steps gives the same
    energies and the same final state.  The step's thermo line carries the
    kinetic energy of the VV step, from before the rescale: rescale_velocities
    does not update e_kin (src/integration.f90:188, src/kinetic.f90:188)."""
    import re

    import oracle_lib as O
    from gpu_util import note, rel
    s = configs.nacl(ncell=2, seed=6, temp_K=300.0)
    nsteps = 8
    FI.write_inputs(tmp_path, s, {1: "Na", 2: "Cl"}, npas=nsteps, extra_control=[" lcsvr = .true."],
                    extra_md=[" nequil = 8", " nequil_period = 1", " temp = 900.0", " taucsvr = 0.01"])
    note("fortran driver csvr equilibration")
    r = _run(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    noise = {int(t): (float(a), float(b))
             for t, a, b in re.findall(r"csvr noise\s+(\d+)\s+(\S+)\s+(\S+)", r.stdout)}
    assert sorted(noise) == list(range(2, nsteps))
    thermo = FI.parse_thermo(r.stdout)
    o = O.oracle(s.desc, s.itype)
    o.set_positions(s.pos)
    o.set_velocities(s.vel)
    o.engforce()
    ci = 0.0
    for it in range(1, nsteps + 1):
        e, u = o.md_vv(1, 0.001 * capi.TIME_UNIT)
        assert abs(thermo["Utot"][it] - u[0]) <= 1e-9 * abs(u[0]), it
        assert abs(thermo["Ekin"][it] - e[0]) <= 1e-8 * e[0], it
        if it in noise:
            ci = o.csvr_rescale(900.0 * capi.BOLTZ_UNIT, 0.01 * capi.TIME_UNIT, *noise[it], ci)
    cont = FI.read_contff(tmp_path / "CONTFF", s.natm)
    assert rel(cont[:, 0:3], o.get_positions()) < 1e-10
    assert rel(cont[:, 3:6], o.get_velocities()) < 1e-8


@pytest.mark.gpu
def test_driver_lautoES_matches_oracle(tmp_path):
    """lautoES (src/coulomb.f90:363-410, dl_poly form): the driver's alpha,
    kmax and clamped cutoff from epsw = 1e-7 give the terms of the oracle run
    at the parameters the same formula gives (configs.auto_ewald)."""
    import copy

    import oracle_lib as O
    from gpu_util import note
    s = configs.nacl(ncell=2, seed=9)
    FI.write_inputs(tmp_path, s, {1: "Na", 2: "Cl"}, npas=0,
                    extra_coulomb=[" lautoES = .true.", " epsw = 1e-7"])
    note("fortran driver lautoES")
    r = _run(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    terms = FI.parse_terms(r.stdout)
    rows = np.array(s.desc.cell[:]).reshape(3, 3)       # lattice vectors
    rc, alpha, kmax = configs.auto_ewald(rows, float(s.desc.cutlongrange), 1e-7)
    assert (alpha, kmax) != (float(s.desc.alphaES), list(s.desc.kES))
    d = copy.deepcopy(s.desc)
    d.alphaES = alpha
    d.kES[:] = kmax
    d.cutlongrange = rc
    o = O.oracle(d, s.itype)
    o.set_positions(s.pos)
    _, to = o.engforce()
    for k in ("u_coul", "u_dir", "u_rec", "u_self"):
        assert abs(terms[k] - to[k]) <= 1e-10 * abs(to[k]), (k, terms[k], to[k])


def _fortran_records(path):
    """Records of a sequential unformatted file (4-byte length markers)."""
    raw = path.read_bytes()
    out, p = [], 0
    while p < len(raw):
        n = int(np.frombuffer(raw[p:p + 4], np.int32)[0])
        out.append(raw[p + 4:p + 4 + n])
        p += 8 + n
    return out


@pytest.mark.gpu
def test_driver_efgall_matches_oracle(tmp_path):
    """doefg: EFGALL as write_EFGALL writes it (src/field.f90:392-448), the
    formatted file (vxx vyy vzz vxy vxz vyz per ion) and the unformatted one
    (the whole efg(3,3,natm), whose lower triangle has no reciprocal part),
    against the oracle's efg_t for the same inputs."""
    import copy

    import oracle_lib as O
    from gpu_util import note, rel
    s = configs.nacl(ncell=2, seed=12)
    d = copy.deepcopy(s.desc)
    d.doefg = 1
    o = O.oracle(d, s.itype)
    o.set_positions(s.pos)
    o.engforce()
    want = o.get_efg()                         # (natm, 3, 3)
    for fmt in (1, 0):
        wd = tmp_path / f"fmt{fmt}"
        wd.mkdir()
        FI.write_inputs(wd, s, {1: "Na", 2: "Cl"}, npas=0, extra_control=[f" iefgall_format = {fmt}"],
                        extra_coulomb=[" doefg = .true."])
        note(f"fortran driver efgall format {fmt}")
        r = _run(wd)
        assert r.returncode == 0, r.stdout + r.stderr
        if fmt:
            rows = np.loadtxt(wd / "EFGALL", skiprows=9, usecols=range(2, 8))
            cols = np.stack([want[:, 0, 0], want[:, 1, 1], want[:, 2, 2], want[:, 0, 1], want[:, 0, 2],
                             want[:, 1, 2]], axis=1)
            assert rel(rows, cols) < 1e-10
        else:
            rec = _fortran_records(wd / "EFGALL")
            assert int(np.frombuffer(rec[0], np.int32)[0]) == s.natm
            g = np.frombuffer(rec[-1], np.float64).reshape(s.natm, 3, 3)   # [i, b, a]
            assert rel(g.transpose(0, 2, 1), want) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("algo", ["scf", "scf_kO_v1", "scf_kO_v2"])
def test_driver_pim_matches_oracle(tmp_path, algo):
    """A polarisable-ion input through the driver's &pimtag (ldip_polar,
    poldip, TT damping tables, SCF controls; src/pim.f90:107-140) with each
    SCF algorithm the library runs, scf_kO_v1 as examples/ex12 uses it: the
    t = 0 energies, u_pol, SCF iteration count and forces against the oracle
    run on the same descriptor."""
    import copy

    import oracle_lib as O
    from gpu_util import note, rel
    s = configs.mgo_pim(cells=(2, 2, 3), a=4.212, seed=9, sigma=0.1, alpha=0.8, rcut=4.2, kmax=8, tol=1e-8)
    FI.write_inputs(tmp_path, s, {1: "Mg", 2: "O"}, npas=0,
                    pim=[f" algo_moment_from_pola = '{algo}'", " omegakO = 0.7"])
    note(f"fortran driver pim {algo}")
    r = _run(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    terms = FI.parse_terms(r.stdout)
    fdrv = np.loadtxt(tmp_path / "FORCES")
    d = copy.deepcopy(s.desc)
    d.calc_md = 1
    d.algo_moment_from_pola = {"scf": 0, "scf_kO_v2": 1, "scf_kO_v1": 2}[algo]
    o = O.oracle(d, s.itype)
    o.set_positions(s.pos)
    fo, to = o.engforce()
    assert int(terms["n_scf"]) == to["n_scf"]
    for k in ("u_bmhft", "u_coul", "u_pol"):
        assert abs(terms[k] - to[k]) <= 1e-10 * abs(to[k]), (k, terms[k], to[k])
    assert rel(fdrv, fo) < 1e-10
