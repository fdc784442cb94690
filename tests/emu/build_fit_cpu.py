"""Builds the CPU emulation of the batched-fit kernels (test infrastructure):
mdff_amd/csrc/fit.hip compiled for the host against tests/emu/hip_cpu_shim.h,
whose hipLaunchKernelGGL runs each launch on the shim's thread-per-lane launcher.  Output:
tests/emu/_build/libfit_cpu.so."""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
OUT = HERE / "_build"


def build() -> Path:
    OUT.mkdir(exist_ok=True)
    src = (ROOT / "mdff_amd" / "csrc" / "fit.hip").read_text()
    subs = [
        ('#include "common.h"', '#include "hip_cpu_shim.h"'),
        ("extern __shared__ double red[];", "double* red = g_shared;"),
    ]
    for a, b in subs:
        if a not in src:
            raise RuntimeError(f"fit.hip changed: {a!r} not found")
        src = src.replace(a, b)
    cpp = OUT / "fit_cpu.cpp"
    cpp.write_text(src)
    lib = OUT / "libfit_cpu.so"
    cmd = ["g++", "-std=c++20", "-O1", "-fPIC", "-shared", "-pthread", f"-I{HERE}", f"-I{ROOT / 'include'}",
           "-o", str(lib), str(cpp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("CPU build of fit.hip failed:\n" + r.stderr[-4000:])
    return lib


if __name__ == "__main__":
    print(build())
