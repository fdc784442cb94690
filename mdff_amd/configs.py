"""Seeded synthetic configurations for the engforce path.

Each builder returns a ``System``: the descriptor (namelist values), 1-based
types, positions, velocities and the MD settings the workload runs with.
Positions follow POSFF's layout (atoms grouped by type, src/read_pos.f90:62-108).
Parameters the workload statement leaves open are fixed here and named in the
docstrings.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .capi import (BOLTZ_UNIT, COUL_UNIT, PRESS_UNIT, TIME_UNIT, MdffDesc, default_desc, set_pair,
                   t3)

AMU_G = 1.660538782e-24  # g per amu (rho_unit, src/constants.f90:48)

MASS = {"Na": 22.98977, "Cl": 35.453, "Si": 28.0855, "O": 15.9994, "Li": 6.941, "F": 18.9984,
        "Mg": 24.305, "A": 1.0}


@dataclass
class System:
    name: str
    desc: MdffDesc
    itype: np.ndarray
    pos: np.ndarray
    vel: np.ndarray
    dt: float = 0.0                   # internal time units
    nsteps: int = 0
    ensemble: str = "nve"
    temp: float = 0.0                 # kB*T in eV
    extra: dict = field(default_factory=dict)

    @property
    def natm(self) -> int:
        return int(self.desc.natm)


def cubic(L: float) -> np.ndarray:
    return np.diag([L, L, L]).astype(np.float64)


def maxwell_boltzmann(rng: np.random.Generator, masses: np.ndarray, temp_K: float) -> np.ndarray:
    """Velocities in internal units (Angstrom per internal time) at temp_K,
    centre-of-mass velocity removed (src/kinetic.f90:414-517)."""
    kT = temp_K * BOLTZ_UNIT
    v = rng.standard_normal((masses.size, 3)) * np.sqrt(kT / masses)[:, None]
    v -= v.mean(axis=0)
    return v


def random_packing(rng: np.random.Generator, n: int, box: np.ndarray, rmin: float) -> np.ndarray:
    """Random sequential addition in an orthorhombic periodic box with a
    minimum separation rmin (cell grid of width >= rmin)."""
    L = np.diag(box).astype(np.float64)
    ncell = np.maximum((L // rmin).astype(int), 1)
    cw = L / ncell
    grid: dict[tuple, list] = {}
    out = np.zeros((n, 3))
    k = 0
    rmin2 = rmin * rmin
    tries = 0
    while k < n:
        tries += 1
        if tries > 200 * n:
            raise RuntimeError("random packing failed: density too high for rmin")
        p = rng.random(3) * L
        c = tuple((p // cw).astype(int) % ncell)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    cc = ((c[0] + dx) % ncell[0], (c[1] + dy) % ncell[1], (c[2] + dz) % ncell[2])
                    for j in grid.get(cc, ()):
                        d = p - out[j]
                        d -= L * np.round(d / L)
                        if d @ d < rmin2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            out[k] = p
            grid.setdefault(c, []).append(k)
            k += 1
    return out


def _base(natm: int, ntype: int, box: np.ndarray, masses: list[float], charges: list[float]) -> MdffDesc:
    d = default_desc(natm, ntype, box)
    for it, (m, q) in enumerate(zip(masses, charges), start=1):
        d.mass[it - 1] = m
        d.qch[it - 1] = q
    return d


def nacl(ncell: int = 4, a: float = 5.64, sigma: float = 0.05, seed: int = 1, temp_K: float = 300.0,
         alpha: float = 0.3, rcut: float = 8.0, kmax: int = 8) -> System:
    """NaCl rock salt, ncell^3 conventional cells (8 ions each), Na +1, Cl -1,
    Born-Huggins-Mayer (Fumi-Tosi) short range + Ewald.  Fumi-Tosi parameters
    (assumed: Tosi & Fumi, J. Phys. Chem. Solids 25, 45 (1964)): b = 0.338e-19 J,
    rho = 0.317 A, sigma_Na = 1.170, sigma_Cl = 1.585 A, Pauling c = 1.25/1.00/0.75,
    C = 1.68/11.2/116 (1e-79 J m^6), D = 0.8/13.9/233 (1e-99 J m^8)."""
    rng = np.random.default_rng(seed)
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.array([[i, j, k] for i in range(ncell) for j in range(ncell) for k in range(ncell)], float)
    na = (cells[:, None, :] + base[None]).reshape(-1, 3) * a
    cl = (cells[:, None, :] + base[None] + np.array([0.5, 0, 0])).reshape(-1, 3) * a
    pos = np.concatenate([na, cl]) + rng.normal(0.0, sigma, (2 * na.shape[0], 3))
    n = pos.shape[0]
    itype = np.array([1] * na.shape[0] + [2] * cl.shape[0], np.int32)
    L = ncell * a
    d = _base(n, 2, cubic(L), [MASS["Na"], MASS["Cl"]], [1.0, -1.0])
    eV = 1.0 / 1.602176487e-19
    b = 0.338e-19 * eV
    rho = 0.317
    sig = {1: 1.170, 2: 1.585}
    cpaul = {(1, 1): 1.25, (1, 2): 1.0, (2, 2): 0.75}
    Cfac = 1e-79 * eV * 1e60
    Dfac = 1e-99 * eV * 1e80
    Cv = {(1, 1): 1.68, (1, 2): 11.2, (2, 2): 116.0}
    Dv = {(1, 1): 0.8, (1, 2): 13.9, (2, 2): 233.0}
    for (i, j), c in cpaul.items():
        set_pair(d, "Abmhftd", i, j, b * c * np.exp((sig[i] + sig[j]) / rho))
        set_pair(d, "Bbmhftd", i, j, 1.0 / rho)
        set_pair(d, "Cbmhftd", i, j, Cv[(i, j)] * Cfac)
        set_pair(d, "Dbmhftd", i, j, Dv[(i, j)] * Dfac)
    d.lbmhft = 1
    d.lcoulomb = 1
    d.cutshortrange = rcut
    d.cutlongrange = rcut
    d.alphaES = alpha
    d.kES[:] = [kmax] * 3
    masses = np.where(itype == 1, MASS["Na"], MASS["Cl"])
    vel = maxwell_boltzmann(rng, masses, temp_K)
    return System("nacl", d, itype, pos, vel, dt=0.001 * TIME_UNIT, nsteps=100, ensemble="nvt-csvr",
                  temp=temp_K * BOLTZ_UNIT, extra={"taucsvr": 0.0})


def sio2(natm: int = 24000, seed: int = 2, density: float = 2.2, rmin: float = 1.4,
         temp_K: float = 300.0, alpha: float = 0.28, rcut: float = 10.0, kmax: int | None = None,
         skin: float = 1.0) -> System:
    """Amorphous SiO2: natm ions (1/3 Si +2.4, 2/3 O -1.2), random packing with
    rmin minimum separation at `density` g/cm3, Buckingham (A, rho, C) + Ewald.
    Buckingham set (assumed): BKS A and rho (van Beest, Kramer, van Santen,
    PRL 64, 1955 (1990)) with the dispersion C divided by 4 so that the random
    packing's close pairs sit on the repulsive wall rather than in the BKS
    short-range collapse.  kmax (assumed) is the smallest n with
    exp(-(pi n / (alpha L))^2) <= 1e-6."""
    assert natm % 3 == 0
    rng = np.random.default_rng(seed)
    nsi = natm // 3
    no = natm - nsi
    mass_tot = nsi * MASS["Si"] + no * MASS["O"]
    L = (mass_tot * AMU_G / density * 1e24) ** (1.0 / 3.0)
    pos = random_packing(rng, natm, cubic(L), rmin)
    itype = np.array([1] * nsi + [2] * no, np.int32)
    d = _base(natm, 2, cubic(L), [MASS["Si"], MASS["O"]], [2.4, -1.2])
    set_pair(d, "Abmhftd", 1, 2, 18003.7572)
    set_pair(d, "Bbmhftd", 1, 2, 4.87318)
    set_pair(d, "Cbmhftd", 1, 2, 133.5381 / 4.0)
    set_pair(d, "Abmhftd", 2, 2, 1388.7730)
    set_pair(d, "Bbmhftd", 2, 2, 2.76000)
    set_pair(d, "Cbmhftd", 2, 2, 175.0000 / 4.0)
    d.lbmhft = 1
    d.lcoulomb = 1
    d.cutshortrange = rcut
    d.cutlongrange = rcut
    d.skindiff = skin
    d.alphaES = alpha
    if kmax is None:
        kmax = int(np.ceil(alpha * L * np.sqrt(np.log(1e6)) / np.pi))
    d.kES[:] = [kmax] * 3
    masses = np.where(itype == 1, MASS["Si"], MASS["O"])
    vel = maxwell_boltzmann(rng, masses, temp_K)
    return System("sio2", d, itype, pos, vel, dt=0.001 * TIME_UNIT, nsteps=1000, ensemble="nve")


def lif(natm: int = 200000, seed: int = 3, density: float = 1.80, temp_K: float = 1200.0,
        alpha: float = 0.35, rcut: float = 9.0, kmax: int | None = None, skin: float = 1.0,
        rmin: float = 1.5) -> System:
    """LiF-like molten salt: natm ions (Li +1, F -1) in a cubic box at `density`
    (assumed 1.80 g/cm3), seeded random positions (rmin separation), Buckingham
    + Ewald.  Buckingham (assumed, Catlow-type): Li-F A=426.480 rho=0.2766,
    F-F A=1127.7 rho=0.2753 C=15.83, Li-Li A=0."""
    assert natm % 2 == 0
    rng = np.random.default_rng(seed)
    nh = natm // 2
    L = (nh * (MASS["Li"] + MASS["F"]) * AMU_G / density * 1e24) ** (1.0 / 3.0)
    pos = random_packing(rng, natm, cubic(L), rmin)
    itype = np.array([1] * nh + [2] * nh, np.int32)
    d = _base(natm, 2, cubic(L), [MASS["Li"], MASS["F"]], [1.0, -1.0])
    set_pair(d, "Abmhftd", 1, 2, 426.480)
    set_pair(d, "Bbmhftd", 1, 2, 1.0 / 0.2766)
    set_pair(d, "Abmhftd", 2, 2, 1127.7)
    set_pair(d, "Bbmhftd", 2, 2, 1.0 / 0.2753)
    set_pair(d, "Cbmhftd", 2, 2, 15.83)
    d.lbmhft = 1
    d.lcoulomb = 1
    d.cutshortrange = rcut
    d.cutlongrange = rcut
    d.skindiff = skin
    d.alphaES = alpha
    if kmax is None:
        kmax = int(np.ceil(alpha * L * np.sqrt(np.log(1e6)) / np.pi))
    d.kES[:] = [kmax] * 3
    masses = np.where(itype == 1, MASS["Li"], MASS["F"])
    vel = maxwell_boltzmann(rng, masses, temp_K)
    return System("lif", d, itype, pos, vel, dt=0.001 * TIME_UNIT, nsteps=10000, ensemble="nvt-nhcn",
                  temp=temp_K * BOLTZ_UNIT, extra={"nhc_n": 3, "timesca_thermo": 0.1 * TIME_UNIT})


def mgo_pim(cells: tuple[int, int, int] = (10, 10, 15), a: float = 4.212, seed: int = 4,
            sigma: float = 0.05, temp_K: float = 300.0, alpha: float = 0.3, rcut: float = 10.0,
            kmax: int | None = None, pol: float = 1.2, tt_b: float = 3.0, tt_c: float = 1.0,
            tt_k: int = 4, tol: float = 1e-7, skin: float = 1.0) -> System:
    """Polarisable MgO-like oxide: rock salt (Mg +2, O -2) with seeded Gaussian
    displacements, induced dipoles on O (isotropic polarisability `pol` A^3),
    Tang-Toennies damping of the charge->dipole field (b, c, order), Buckingham
    short range (assumed, Lewis & Catlow: Mg-O A=821.6 rho=0.3242;
    O-O A=22764.0 rho=0.149 C=27.88), SCF tolerance `tol`.  10x10x15 cells gives
    12,000 ions at 3.58 g/cm3."""
    rng = np.random.default_rng(seed)
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    nx, ny, nz = cells
    cl = np.array([[i, j, k] for i in range(nx) for j in range(ny) for k in range(nz)], float)
    mg = (cl[:, None, :] + base[None]).reshape(-1, 3) * a
    ox = (cl[:, None, :] + base[None] + np.array([0.5, 0, 0])).reshape(-1, 3) * a
    pos = np.concatenate([mg, ox]) + rng.normal(0.0, sigma, (2 * mg.shape[0], 3))
    n = pos.shape[0]
    itype = np.array([1] * mg.shape[0] + [2] * ox.shape[0], np.int32)
    box = np.diag([nx * a, ny * a, nz * a])
    d = _base(n, 2, box, [MASS["Mg"], MASS["O"]], [2.0, -2.0])
    set_pair(d, "Abmhftd", 1, 2, 821.6)
    set_pair(d, "Bbmhftd", 1, 2, 1.0 / 0.3242)
    set_pair(d, "Abmhftd", 2, 2, 22764.0)
    set_pair(d, "Bbmhftd", 2, 2, 1.0 / 0.149)
    set_pair(d, "Cbmhftd", 2, 2, 27.88)
    d.lbmhft = 1
    d.lcoulomb = 1
    d.cutshortrange = rcut
    d.cutlongrange = rcut
    d.skindiff = skin
    d.alphaES = alpha
    if kmax is None:
        Lmax = max(nx, ny, nz) * a
        kmax = int(np.ceil(alpha * Lmax * np.sqrt(np.log(1e6)) / np.pi))
    d.kES[:] = [kmax] * 3
    set_pim(d, polar_type=2, pol=pol, tt_b=tt_b, tt_c=tt_c, tt_k=tt_k, ntype=2)
    d.conv_tol_ind = tol
    masses = np.where(itype == 1, MASS["Mg"], MASS["O"])
    vel = maxwell_boltzmann(rng, masses, temp_K)
    return System("mgo_pim", d, itype, pos, vel, dt=0.001 * TIME_UNIT, nsteps=500, ensemble="nvt-nhcn",
                  temp=temp_K * BOLTZ_UNIT, extra={"nhc_n": 3, "timesca_thermo": 0.1 * TIME_UNIT})


def set_pim(d: MdffDesc, polar_type: int, pol: float, tt_b: float, tt_c: float, tt_k: int,
            ntype: int) -> None:
    """Isotropic polarisability on one type and TT damping of the field on it
    from every charged type: ldip_damping(p, p, j) = T for all j
    (src/pim.f90:78-81; symmetric in the last two indices, :186-197)."""
    p = polar_type
    d.ldip_polar[p - 1] = 1
    for a in range(3):
        d.poldip[(p - 1) + 16 * a + 48 * a] = pol
    for j in range(1, ntype + 1):
        for (i, jj) in ((p, j), (j, p)):
            k = t3(p, i, jj)
            d.ldip_damping[k] = 1
            d.pol_damp_b[k] = tt_b
            d.pol_damp_c[k] = tt_c
            d.pol_damp_k[k] = tt_k


def two_charges() -> System:
    """examples/ex04 POSFF.twocharges: +1/-1 at x = -0.5/+0.5 in a 10^3 box,
    reduced units, alpha 0.8, kmax 11, cutlongrange 5 (control_twocharges.F)."""
    d = _base(2, 2, cubic(10.0), [1.0, 1.0], [1.0, -1.0])
    d.coul_unit = 1.0
    d.press_unit = 1.0
    d.lcoulomb = 1
    d.cutlongrange = 5.0
    d.alphaES = 0.8
    d.kES[:] = [11, 11, 11]
    pos = np.array([[-0.5, 0.0, 0.0], [0.5, 0.0, 0.0]])
    return System("two_charges", d, np.array([1, 2], np.int32), pos, np.zeros((2, 3)))


def two_dipoles() -> System:
    """examples/ex04 POSFF.twodipoles: static dipoles (1,0,0) on both ions."""
    s = two_charges()
    d = s.desc
    d.qch[0] = 0.0
    d.qch[1] = 0.0
    d.dip[0 + 16 * 0] = 1.0
    d.dip[1 + 16 * 0] = 1.0
    s.name = "two_dipoles"
    return s


def lj_cluster(xyz: np.ndarray) -> System:
    """examples/ex02: LJ cluster in a 100^3 box, eps = sigma = 1, rc = 100,
    ctrunc = 'linear', reduced units (examples/ex02/config/control.F)."""
    n = xyz.shape[0]
    d = _base(n, 1, cubic(100.0), [1.0], [0.0])
    d.lnmlj = 1
    d.cutshortrange = 100.0
    d.trunc = 1
    set_pair(d, "epslj", 1, 1, 1.0)
    set_pair(d, "sigmalj", 1, 1, 1.0)
    d.coul_unit = 1.0
    d.press_unit = 1.0
    return System("lj_cluster", d, np.ones(n, np.int32), np.asarray(xyz, float), np.zeros((n, 3)))


def fcc_lj(ncell: int = 10, L: float = 15.414361, rc: float = 6.0) -> System:
    """The 4000-atom FCC LJ sample run of doc/mdff.tex:270-421: 10^3 FCC cells
    in a cubic box of side 15.414361, eps = sigma = 1, rc = 6, linear shift."""
    a = L / ncell
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.array([[i, j, k] for i in range(ncell) for j in range(ncell) for k in range(ncell)], float)
    pos = (cells[:, None, :] + base[None]).reshape(-1, 3) * a
    n = pos.shape[0]
    d = _base(n, 1, cubic(L), [1.0], [0.0])
    d.lnmlj = 1
    d.cutshortrange = rc
    d.trunc = 1
    set_pair(d, "epslj", 1, 1, 1.0)
    set_pair(d, "sigmalj", 1, 1, 1.0)
    return System("fcc_lj", d, np.ones(n, np.int32), pos, np.zeros((n, 3)))


def scaled(builder, natm: int, **kw) -> System:
    """A builder's workload at another size (same density and settings)."""
    return builder(natm=natm, **kw)


__all__ = ["System", "nacl", "sio2", "lif", "mgo_pim", "two_charges", "two_dipoles", "lj_cluster",
           "fcc_lj", "maxwell_boltzmann", "random_packing", "set_pim", "COUL_UNIT", "PRESS_UNIT"]


def relax_sd(engine, pos: np.ndarray, iters: int = 150, max_step: float = 0.05,
             gain: float = 0.02) -> tuple[np.ndarray, float]:
    """Steepest descent on the engine's engforce: every iteration moves all
    ions along F by min(max_step, gain*max|F|) scaled by F/max|F|.  Used to
    take a random packing off its overlapping contacts before dynamics (a raw
    packing releases eV per ion and heats to ~1e5 K).  After every move the
    engine's Verlet lists are checked (vnlistcheck, as the MD loop does each
    step): engforce itself reuses the lists.  Returns the relaxed positions
    and the final max |F|."""
    p = np.array(pos, dtype=np.float64, copy=True)
    fmax = float("inf")
    for _ in range(iters):
        engine.set_positions(p)
        engine.vnlistcheck()
        f, _ = engine.engforce()
        fmax = float(np.abs(f).max())
        if fmax == 0.0:
            break
        p += min(max_step, gain * fmax) * f / fmax
    engine.set_positions(p)
    engine.vnlistcheck()
    return p, fmax


def auto_ewald(cell: np.ndarray, rc: float, epsw: float = 1e-6) -> tuple[float, float, list[int]]:
    """Automatic Ewald parameters of the reference (ewald_param with lautoES,
    src/coulomb.f90:363-410, dl_poly form): the real-space cutoff clamped to
    half the smallest cell width, alpha and kmax from the accuracy epsw.
    cell rows are the lattice vectors.  Returns (rc, alpha, kmax)."""
    A = np.asarray(cell, float)
    omega = abs(np.linalg.det(A))
    widths = [omega / np.linalg.norm(np.cross(A[(k + 1) % 3], A[(k + 2) % 3])) for k in range(3)]
    rc = min(rc, 0.5 * min(widths))
    eps = min(abs(epsw), 0.5)
    tol = np.sqrt(abs(np.log(eps * rc)))
    alpha = np.sqrt(abs(np.log(eps * rc * tol))) / rc
    tol1 = np.sqrt(-np.log(eps * rc * (2.0 * tol * alpha) ** 2))
    kmax = [int(np.rint(0.25 + np.linalg.norm(A[k]) * alpha * tol1 / np.pi)) for k in range(3)]
    return float(rc), float(alpha), kmax


# hidden "true" parameters of the fit workload: MgO-like PIM oxide
FIT_THETA_STAR = {
    ("A", 1, 2): 821.6, ("RHO", 1, 2): 0.3242,
    ("A", 2, 2): 22764.0, ("RHO", 2, 2): 0.149, ("C", 2, 2): 27.88,
    ("ALPHA", 2, 0): 1.2,
    ("TTB", 2, 1): 3.0, ("TTC", 2, 1): 1.0, ("TTB", 2, 2): 3.0, ("TTC", 2, 2): 1.0,
}


def fit_snapshots(nsnap: int = 256, nmin: int = 96, nmax: int = 1536, seed: int = 5, a: float = 4.212,
                  sigma: float = 0.08, rc: float = 8.0, epsw: float = 1e-6, tol: float = 1e-7) -> list[System]:
    """Snapshots of the batched fit: rock-salt MgO supercells (n1 x n2 x n3
    conventional cells, 8 n1 n2 n3 ions, sizes log-uniform in [nmin, nmax],
    aspect ratio <= 2) with seeded Gaussian displacements sigma, each with its
    own Ewald parameters from auto_ewald (cutoff min(rc, half width)).  The
    force field is mgo_pim's at FIT_THETA_STAR (Buckingham + PIM on O with
    TT damping of order 4)."""
    rng = np.random.default_rng(seed)
    shapes = sorted({tuple(sorted((i, j, k))) for i in range(2, 13) for j in range(2, 13) for k in range(2, 13)
                     if nmin <= 8 * i * j * k <= nmax and max(i, j, k) <= 2 * min(i, j, k)})
    sizes = np.array([8 * s[0] * s[1] * s[2] for s in shapes])
    out = []
    for s in range(nsnap):
        target = np.exp(rng.uniform(np.log(nmin), np.log(nmax)))
        cand = np.flatnonzero(np.abs(np.log(sizes / target)) == np.abs(np.log(sizes / target)).min())
        cells = list(shapes[int(cand[0])])
        rng.shuffle(cells)
        base = mgo_pim(cells=tuple(cells), a=a, seed=int(rng.integers(1 << 30)), sigma=sigma, tol=tol)
        d = base.desc
        box = np.asarray(d.cell[:], float).reshape(3, 3)
        rcs, alpha, kmax = auto_ewald(box, rc, epsw)
        d.cutshortrange = rcs
        d.cutlongrange = rcs
        d.alphaES = alpha
        d.kES[:] = kmax
        d.skindiff = 0.1
        base.name = f"fit{s}"
        base.extra.update({"cells": tuple(cells), "rc": rcs, "alpha": alpha, "kmax": kmax})
        out.append(base)
    return out
