"""Short run of the a-SiO2 24,000-ion workload (the bench configuration) for
profiling with rocprofv3: steepest-descent relaxation, then `--steps`
velocity-Verlet steps of one precision form, then a device-to-device copy of
`--copy-mb` MiB (hipMemcpy device to device) whose counter readings
calibrate FETCH_SIZE / WRITE_SIZE (the copy reads and writes exactly that
many bytes).  `--fit N` instead runs N forward + gradient evaluations of
the batched fit (256 PIM oxide snapshots of 96-1536 ions, ten parameters,
the bench's fit leg without its autotuner: blocks of 256 atoms or
k-vectors).

    python -m mdff_amd.profile_workload --steps 20 --precision mixed
    python -m mdff_amd.profile_workload --fit 1 --precision mixed
"""
from __future__ import annotations

import argparse
import copy
import sys

from . import capi, configs


def fit_workload(args) -> int:
    import numpy as np
    from . import fit as F
    params = [("A", 1, 2), ("RHO", 1, 2), ("A", 2, 2), ("RHO", 2, 2), ("C", 2, 2), ("ALPHA", 2, 0),
              ("TTB", 2, 1), ("TTC", 2, 1), ("TTB", 2, 2), ("TTC", 2, 2)]
    snaps = configs.fit_snapshots(256)
    theta = F.theta_of(params, configs.FIT_THETA_STAR)
    natm = sum(s.natm for s in snaps)
    P = F.FitProblem(snaps, params)
    P.set_targets(np.zeros((natm, 3)), np.zeros((len(snaps), 9)), 0.05)
    P.set_launch(256, 1)
    prec = capi.PREC_MIXED if args.precision == "mixed" else capi.PREC_FP64
    for _ in range(args.fit):
        P.eval(theta, prec, grad=True)
    P.close()
    print(f"profile_workload: {args.fit} {args.precision} fit evaluations on {len(snaps)} snapshots done", flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--natm", type=int, default=24000)
    ap.add_argument("--relax", type=int, default=150)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--copy-mb", type=int, default=256)
    ap.add_argument("--fit", type=int, default=0, help="batched-fit evaluations instead of MD steps")
    args = ap.parse_args(argv)
    if args.fit > 0:
        return fit_workload(args)
    s = configs.sio2(natm=args.natm)
    if args.relax > 0:
        eng = capi.gpu_engine(copy.deepcopy(s.desc), s.itype)
        s.pos, _ = configs.relax_sd(eng, s.pos, iters=args.relax)
        eng.close()
    d = copy.deepcopy(s.desc)
    d.precision = capi.PREC_MIXED if args.precision == "mixed" else capi.PREC_FP64
    eng = capi.gpu_engine(d, s.itype)
    eng.set_positions(s.pos)
    eng.set_velocities(s.vel)
    eng.engforce()
    eng.md_vv(args.steps, s.dt)
    eng.close()
    print(f"profile_workload: {args.steps} {args.precision} VV steps on {s.natm} ions done", flush=True)
    if args.copy_mb > 0:
        # device-to-device copy through the HIP runtime (no framework in the
        # profiled process): reads and writes copy_mb MiB once each
        import ctypes as C
        hip = C.CDLL("libamdhip64.so")
        nbytes = args.copy_mb * (1 << 20)
        src, dst = C.c_void_p(), C.c_void_p()
        if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
            raise RuntimeError("hipMalloc failed")
        hip.hipMemset(src, 1, C.c_size_t(nbytes))
        hip.hipDeviceSynchronize()
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3):       # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
        hip.hipDeviceSynchronize()
        hip.hipFree(src)
        hip.hipFree(dst)
        print(f"profile_workload: calibration copy of {args.copy_mb} MiB done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
