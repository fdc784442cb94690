"""Reference-format inputs (namelist control file + POSFF) for the Fortran
driver, written from a System of mdff_amd.configs.  Reals are written with
repr(), so the driver reads back the same doubles."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from mdff_amd import capi

ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "mdff_amd" / "lib" / "mdff_gpu_driver"


def _r(x) -> str:
    return repr(float(x))


def write_inputs(workdir: Path, s, names: dict[int, str], npas: int = 0, dt_ps: float = 0.001,
                 extra_control: list[str] | None = None) -> Path:
    """control.F and POSFF (Cartesian, rvn) for a Buckingham/BMH + Ewald system."""
    d = s.desc
    ntype = int(d.ntype)
    L = [
        "&controltag",
        f" lbmhft = {'.true.' if d.lbmhft else '.false.'}",
        f" lbmhftd = {'.true.' if d.lbmhftd else '.false.'}",
        f" lcoulomb = {'.true.' if d.lcoulomb else '.false.'}",
        f" lvnlist = {'.true.' if d.lvnlist else '.false.'}",
        f" cutshortrange = {_r(d.cutshortrange)}",
        f" cutlongrange = {_r(d.cutlongrange)}",
        f" skindiff = {_r(d.skindiff)}",
        " posff_data = 'rvn'",
        *(extra_control or []),
        "&end",
        "&mdtag",
        " integrator = 'nve-vv'",
        f" npas = {npas}",
        f" dt = {_r(dt_ps)}",
        " nprint = 1",
        "&end",
        "&non_bondedtag",
    ]
    for f in ("Abmhftd", "Bbmhftd", "Cbmhftd", "Dbmhftd", "BDbmhftd"):
        arr = getattr(d, f)
        for i in range(1, ntype + 1):
            for j in range(i, ntype + 1):
                L.append(f" {f}({i},{j}) = {_r(arr[capi.t2(i, j)])}")
    L += [
        "&end",
        "&coulombtag",
        f" alphaES = {_r(d.alphaES)}",
        f" kES = {d.kES[0]}, {d.kES[1]}, {d.kES[2]}",
        "&end",
        "&fieldtag",
        " mass = " + ", ".join(_r(d.mass[k]) for k in range(ntype)),
        " qch = " + ", ".join(_r(d.qch[k]) for k in range(ntype)),
        "&end",
    ]
    workdir.mkdir(parents=True, exist_ok=True)
    (workdir / "control.F").write_text("\n".join(L) + "\n")
    cell = np.asarray(d.cell[:], float).reshape(3, 3)   # rows = lattice vectors
    with open(workdir / "POSFF", "w") as f:
        f.write(f"{s.natm}\n{s.name}\n")
        for r in range(3):
            f.write(" ".join(_r(x) for x in cell[r]) + "\n")
        f.write(f"{ntype}\n" + " ".join(names[k] for k in range(1, ntype + 1)) + "\n")
        f.write(" ".join(str(int((s.itype == k).sum())) for k in range(1, ntype + 1)) + "\nCartesian\n")
        for i in range(s.natm):
            p, v = s.pos[i], s.vel[i]
            f.write(f"{names[int(s.itype[i])]} {_r(p[0])} {_r(p[1])} {_r(p[2])} "
                    f"{_r(v[0])} {_r(v[1])} {_r(v[2])}\n")
    return workdir / "control.F"


def parse_terms(stdout: str) -> dict:
    out = {}
    for line in stdout.splitlines():
        if "=" not in line:
            continue
        k, v = line.split("=", 1)
        k = k.strip()
        vals = v.split()
        try:
            nums = [float(x) for x in vals]
        except ValueError:
            continue
        out[k] = nums[0] if len(nums) == 1 else np.array(nums)
    return out


def parse_thermo(stdout: str) -> np.ndarray:
    rows = []
    seen = False
    for line in stdout.splitlines():
        if line.strip().startswith("step"):
            seen = True
            continue
        if seen:
            parts = line.split()
            if len(parts) == 6:
                rows.append([float(x) for x in parts])
    return np.array(rows)


def read_contff(path: Path, natm: int) -> np.ndarray:
    lines = path.read_text().splitlines()
    data = np.array([[float(x) for x in ln.split()[1:]] for ln in lines[9:9 + natm]])
    return data   # natm x 9: r, v, f
