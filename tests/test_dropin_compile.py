"""The drop-in MODULE engforce_driver (mdff_amd/fortran/engforce_driver_gpu.f90)
compiles against the reference's own module globals: the module interfaces
of the reference's engforce-side modules are generated from its sources
(make -C oracle refmods: flang -fsyntax-only, nothing linked), then the
interface module and the drop-in are compiled to objects with them.  Needs
the reference tree and flang (this container); skipped elsewhere."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference/src")
FORT = ROOT / "mdff_amd" / "fortran"


@pytest.mark.skipif(not REF.exists() or shutil.which("flang") is None,
                    reason="needs the reference sources and flang")
def test_dropin_engforce_driver_compiles_against_reference_modules(tmp_path):
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "refmods"], check=True, capture_output=True)
    mods = ROOT / "oracle" / "_ref" / "mods"
    assert (mods / "config.mod").exists() and (mods / "thermodynamic.mod").exists()
    out = tmp_path / "mods"
    out.mkdir()
    for src in ("mdff_gpu_iface.f90", "engforce_driver_gpu.f90"):
        r = subprocess.run(["flang", "-cpp", "-I", str(mods), "-I", str(out), "-module-dir", str(out), "-c",
                            str(FORT / src), "-o", str(tmp_path / (src + ".o"))],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
    assert (out / "engforce_driver.mod").exists()
