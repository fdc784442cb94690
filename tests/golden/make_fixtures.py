"""Writes the golden fixtures under tests/golden/ from data the reference
holds (run in the build container, where /root/reference exists).

  ccd_lj.json     : Cambridge Cluster Database LJ geometries and energies
                    shipped by the reference (examples/ex02/config/<N>,
                    examples/ex02/config/REFERENCE).
  ref_scalar.json : outputs of the reference's own errfc (src/functions.f90:28-40)
                    and TT_damping_functions (src/tt_damp.f90:68-97), compiled
                    from its sources into oracle/_ref (make -C oracle ref).
  ex04.json       : Nymand & Linse Ewald values quoted by the reference
                    (examples/ex04/run_ex4.sh:145,179).
"""
from __future__ import annotations

import ctypes as C
import json
import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
REF = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")


def ccd():
    cfg = REF / "examples" / "ex02" / "config"
    energies = {}
    for line in (cfg / "REFERENCE").read_text().splitlines():
        if line.startswith("#") or not line.strip():
            continue
        k, v = line.split()[:2]
        energies[k] = float(v)
    out = {}
    for name in ["5", "13", "19", "26", "38", "38i", "55", "75i", "100", "147"]:
        xyz = np.loadtxt(cfg / name).reshape(-1, 3)
        out[name] = {"energy": energies[name], "xyz": xyz.tolist()}
    return out


def ref_scalar():
    lib = C.CDLL(str(ROOT / "oracle" / "_ref" / "libmdff_ref.so"))
    lib.ref_errfc.argtypes = [C.c_double]
    lib.ref_errfc.restype = C.c_double
    dp = C.POINTER(C.c_double)
    lib.ref_tt_damp.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, dp, dp]
    xs = np.concatenate([np.linspace(-4.0, 8.0, 241), [0.0, 1e-8, 0.3 * 8.0, 0.28 * 10.0, 2.0 ** -20]])
    erfc = [float.hex(lib.ref_errfc(float(x))) for x in xs]
    tt = []
    f, fd = C.c_double(), C.c_double()
    for b, c, order in [(3.0, 1.0, 4), (3.0, 1.0, 6), (2.0, 0.7, 8), (3.5, 1.2, 1), (1.7, 1.0, 8)]:
        for r in np.linspace(0.5, 10.0, 39):
            lib.ref_tt_damp(b, c, float(r), order, C.byref(f), C.byref(fd))
            tt.append([b, c, order, float(r), float.hex(f.value), float.hex(fd.value)])
    return {"errfc_x": [float.hex(float(x)) for x in xs], "errfc": erfc, "tt": tt}


def main():
    (HERE / "ccd_lj.json").write_text(json.dumps(ccd()))
    (HERE / "ref_scalar.json").write_text(json.dumps(ref_scalar()))
    (HERE / "ex04.json").write_text(json.dumps({
        "source": "examples/ex04/run_ex4.sh:140-216 (Nymand & Linse, J. Chem. Phys. 112, 6152 (2000))",
        "two_charges": {"U": -1.0021255, "E1x": 0.9956865, "V1xx": 2.0003749, "V1yy": -1.0001874,
                        "f1x": 0.9956865},
        "two_dipoles": {"U": -2.0087525, "E1x": 2.0087525, "V1xx": 5.9992460, "V1yy": -2.9996230,
                        "f1x": 5.9992460},
    }, indent=1))


if __name__ == "__main__":
    main()
