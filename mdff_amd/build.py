"""Builds the gfx950 library in-tree: mdff_amd/lib/libmdff_gpu.so.

Every HIP source is compiled by hipcc for gfx950 only, with packed FP32 ops
disabled (-packed-fp32-ops).  nlist.hip is also compiled with
-ffp-contract=off: its r^2 tests must round exactly like the reference.
The Fortran interface module and the Fortran driver are compiled with the
ROCm flang when it is present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "csrc"
OUT = HERE / "lib"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")

COMMON = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
PER_FILE = {"nlist.hip": ["-ffp-contract=off"]}
SOURCES = ["nlist.hip", "realspace.hip", "recip.hip", "mdff_gpu.hip", "fit.hip", "noalloc.hip", "efg.hip"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    noisy = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if noisy:
        sys.stderr.write("\n".join(noisy) + "\n")


def build_hip(verbose: bool = False) -> Path:
    OUT.mkdir(exist_ok=True)
    objdir = OUT / "obj"
    objdir.mkdir(exist_ok=True)

    def one(src: str) -> Path:
        obj = objdir / (src + ".o")
        cmd = [HIPCC, *COMMON, *PER_FILE.get(src, []), "-c", str(SRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return obj

    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(one, SOURCES))
    lib = OUT / "libmdff_gpu.so"
    tmp = OUT / "libmdff_gpu.so.tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, lib)
    return lib


def build_fortran(verbose: bool = False) -> Path | None:
    """Fortran iso_c_binding module + the Fortran driver program (host side
    in the reference's language)."""
    fc = shutil.which("flang") or str(ROCM / "llvm" / "bin" / "flang")
    fdir = HERE / "fortran"
    if not Path(fc).exists() or not (fdir / "mdff_gpu_iface.f90").exists():
        return None
    mod = OUT / "fmod"
    mod.mkdir(parents=True, exist_ok=True)
    srcs = ["mdff_gpu_iface.f90", "mdff_gpu_driver.f90"]
    objs = []
    for s in srcs:
        if not (fdir / s).exists():
            continue
        o = mod / (s + ".o")
        cmd = [fc, "-O2", "-fPIC", "-J", str(mod), "-c", str(fdir / s), "-o", str(o)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        objs.append(o)
    exe = OUT / "mdff_gpu_driver"
    if len(objs) == 2:
        _run([fc, "-o", str(exe), *map(str, objs), "-L", str(OUT), "-lmdff_gpu",
              f"-Wl,-rpath,{OUT}", "-Wl,-rpath,$ORIGIN"])
    return exe


def main(argv: list[str]) -> None:
    v = "-v" in argv
    lib = build_hip(verbose=v)
    print(f"built {lib}")
    exe = build_fortran(verbose=v)
    if exe:
        print(f"built {exe}")


if __name__ == "__main__":
    main(sys.argv[1:])
