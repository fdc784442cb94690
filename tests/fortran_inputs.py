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
                 extra_control: list[str] | None = None, extra_md: list[str] | None = None,
                 extra_coulomb: list[str] | None = None, pim: list[str] | None = None) -> Path:
    """control.F and POSFF (Cartesian, rvn) for a Buckingham/BMH + Ewald system;
    pim: extra &pimtag lines added to the polarisabilities and TT damping the
    descriptor holds (written only when some type is polarisable)."""
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
        *(extra_md or []),
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
        *(extra_coulomb or []),
        "&end",
        "&fieldtag",
        " mass = " + ", ".join(_r(d.mass[k]) for k in range(ntype)),
        " qch = " + ", ".join(_r(d.qch[k]) for k in range(ntype)),
        "&end",
    ]
    if any(d.ldip_polar[k] for k in range(ntype)):
        L.append("&pimtag")
        for it in range(1, ntype + 1):
            if not d.ldip_polar[it - 1]:
                continue
            L.append(f" ldip_polar({it}) = .true.")
            for a in range(1, 4):
                for b in range(1, 4):
                    v = d.poldip[(it - 1) + 16 * (a - 1) + 48 * (b - 1)]
                    if v != 0.0:
                        L.append(f" poldip({it},{a},{b}) = {_r(v)}")
        for k in range(1, ntype + 1):
            for i in range(1, ntype + 1):
                for j in range(1, ntype + 1):
                    x = capi.t3(k, i, j)
                    if d.ldip_damping[x]:
                        L += [f" ldip_damping({k},{i},{j}) = .true.", f" pol_damp_b({k},{i},{j}) = {_r(d.pol_damp_b[x])}",
                              f" pol_damp_c({k},{i},{j}) = {_r(d.pol_damp_c[x])}",
                              f" pol_damp_k({k},{i},{j}) = {d.pol_damp_k[x]}"]
        L += [f" conv_tol_ind = {_r(d.conv_tol_ind)}", f" min_scf_pol_iter = {d.min_scf_pol_iter}",
              f" max_scf_pol_iter = {d.max_scf_pol_iter}", f" extrapolate_order = {d.extrapolate_order}",
              *(pim or []), "&end"]
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


THERMO_KEYS = ["step", "time", "Ekin", "Temp", "Utot", "U_vdw", "U_coul", "Pressure", "Pvir_vdw",
               "Pvir_coul", "Pvir_tot", "volume", "a cell", "b cell", "c cell", "density", "Etot", "Htot"]


def parse_thermo(stdout: str) -> dict:
    """The 'Thermodynamic information' blocks of the standard output
    (write_thermo 'std', src/thermo.f90:395-420): one array per quantity."""
    out = {k: [] for k in THERMO_KEYS}
    for line in stdout.splitlines():
        if "=" not in line:
            continue
        k, v = line.split("=", 1)
        k = k.strip()
        if k in out:
            out[k].append(float(v))
    return {k: np.array(v) for k, v in out.items()}


def parse_osziff(path: Path) -> np.ndarray:
    """OSZIFF (write_thermo 'osz', src/thermo.f90:367-383): pairs of lines
    (I9,2X,6E16.8) under a dashed header every 10 records; rows of
    step, time, Etot, Ekin, Utot, U_vdw, U_coul, Temp, Pressure, Pvir_vdw,
    Pvir_coul, Omega, Htot."""
    recs = []
    lines = [ln for ln in path.read_text().splitlines() if ln.strip() and not ln.startswith("-")
             and "step" not in ln]
    for a, b in zip(lines[0::2], lines[1::2]):
        ia, ib = int(a[:9]), int(b[:9])
        assert ia == ib
        fa = [float(a[11 + 16 * j:11 + 16 * (j + 1)]) for j in range(6)]
        fb = [float(b[11 + 16 * j:11 + 16 * (j + 1)]) for j in range(6)]
        recs.append([ia] + fa + fb)
    return np.array(recs)


def read_contff(path: Path, natm: int) -> np.ndarray:
    lines = path.read_text().splitlines()
    data = np.array([[float(x) for x in ln.split()[1:]] for ln in lines[9:9 + natm]])
    return data   # natm x 9: r, v, f


def read_trajff(path: Path, natm: int, ncol: int) -> list[np.ndarray]:
    """Formatted TRAJFF frames (write_trajff_xyz, src/config.f90:565-595):
    9 header lines, then natm lines of a type label and ncol reals."""
    lines = path.read_text().splitlines()
    frames = []
    k = 0
    while k + 9 <= len(lines):
        n = int(lines[k])
        assert n == natm
        rows = lines[k + 9:k + 9 + natm]
        frames.append(np.array([[float(x) for x in ln.split()[1:1 + ncol]] for ln in rows]))
        k += 9 + natm
    return frames
