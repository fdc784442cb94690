"""Builds the CPU emulation of the batched-fit kernels (test infrastructure):
mdff_amd/csrc/fit.hip compiled for the host against tests/emu/hip_cpu_shim.h,
with the launches redirected to the shim's thread-per-lane launcher.  Output:
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
        *[(f"hipLaunchKernelGGL((k_fit<{t}, {lb}>), dim3(f->grid), dim3(f->block), shm, f->stream, a);",
           f"emu_launch((k_fit<{t}, {lb}>), dim3(f->grid), dim3(f->block), a);")
          for t in ("float", "double") for lb in (256, 1024)],
    ]
    for a, b in subs:
        if a not in src:
            raise RuntimeError(f"fit.hip changed: {a!r} not found")
        src = src.replace(a, b)
    # the table kernels: plain function calls through the launcher
    i0 = src.index("  hipLaunchKernelGGL((k_fit_tables<float>)")
    i1 = src.index("  FIT_TRY(f, hipGetLastError());", i0)
    src = src[:i0] + """  { struct TA { int ns; const FitSnap* sn; const double *x, *y, *z; void* t; };
    emu_launch([](TA q) { k_fit_tables<float>(q.ns, q.sn, q.x, q.y, q.z, q.t); }, dim3(tg), dim3(8),
               TA{nsnap, f->d_snaps, f->d_sx, f->d_sy, f->d_sz, f->d_tab32});
    emu_launch([](TA q) { k_fit_tables<double>(q.ns, q.sn, q.x, q.y, q.z, q.t); }, dim3(tg), dim3(8),
               TA{nsnap, f->d_snaps, f->d_sx, f->d_sy, f->d_sz, f->d_tab64}); }
""" + src[i1:]
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
