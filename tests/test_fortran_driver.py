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
    assert thermo.shape == (nsteps, 6)
    ek, up = [], []
    for _ in range(nsteps):
        e, u = o.md_vv(1, 0.001 * capi.TIME_UNIT)
        ek.append(e[0])
        up.append(u[0])
    assert np.allclose(thermo[:, 4], up, rtol=1e-9, atol=0)
    assert np.allclose(thermo[:, 3], ek, rtol=1e-8, atol=0)
    cont = FI.read_contff(tmp_path / "CONTFF", s.natm)
    assert rel(cont[:, 0:3], o.get_positions()) < 1e-10
    assert rel(cont[:, 3:6], o.get_velocities()) < 1e-8
