"""Rates of the geometry-path kernels on one GPU, arrays resident on the
device: level-set reinitialisation sweeps (ReDist3D_XYZ, LevelSet_3D.f90:
194-261) in FP64 and in FP32 with an int8 sign, one accelerated-diffusion
connectivity run of the reference's length (AccDiff_3D.f90:103-243, 1,500
iterations, averaged) and the component labelling of
both phases (the connectivity mask of RM_UCNCT_3D.f90:79-82), on the
ellipsoid packing with a random-walk tunnel of the 512^3 level-set and
connectivity configuration.  Prints one JSON line.

    python -m elsim_amd.geombench [--n 512] [--sweeps 20] [--report out.json]
"""
import argparse
import ctypes as C
import json
import time

import numpy as np

from ._lib import DeviceArray, check, lib
from .sbm import carve_tunnel, ellipsoid_pack


def timed(fn, reps=1):
    L = lib()
    ms = C.c_float()
    best = None
    for _ in range(reps):
        check(L.elsim_timer_start())
        fn()
        check(L.elsim_timer_stop(C.byref(ms)))
        best = ms.value if best is None else min(best, ms.value)
    return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--sweeps", type=int, default=20)
    ap.add_argument("--ad-iters", type=int, default=1501, help="AD run length (term); the reference runs 1501")
    ap.add_argument("--report", default=None)
    a = ap.parse_args()
    L = lib()
    check(L.elsim_init(0), "elsim_init")
    n = a.n
    t0 = time.perf_counter()
    sdf, nshapes, frac = ellipsoid_pack(n, a.seed, overlap=0.3)
    sdf = carve_tunnel(sdf, a.seed, 3.0)
    t_gen = time.perf_counter() - t0
    N, G = n ** 3, (n + 2) ** 3
    # sign field with mirrored ghosts (as LevelSet_3D builds it after NoGrdBc3D)
    s = np.pad(np.where(sdf > 0, 1.0, -1.0), 1, mode="edge")
    sg64 = DeviceArray.from_host(np.asfortranarray(s).ravel(order="F"))
    sg8 = DeviceArray.from_host(np.asfortranarray(s).ravel(order="F").astype(np.int8))
    f64, f32 = DeviceArray(G, np.float64), DeviceArray(G, np.float32)
    out = {"workload": "geometry_kernels", "grid": [n, n, n], "shapes": nshapes, "solid_fraction": frac,
           "generate_s": t_gen}
    # reinitialisation: lit sweeps from f = 0 (the reference's start), per sweep
    check(L.elsim_memset(f64.ptr, 0, f64.nbytes))
    t = timed(lambda: check(L.elsim_dev_redist3d_f64(a.sweeps, n, n, n, sg64.ptr, f64.ptr, C.c_double(0.1))))
    out["reinit_f64"] = {"us_per_sweep": t * 1e3 / a.sweeps, "bytes_per_voxel": 24,
                         "GBps": 24.0 * G / (t * 1e-3 / a.sweeps) / 1e9}
    check(L.elsim_memset(f32.ptr, 0, f32.nbytes))
    t = timed(lambda: check(L.elsim_dev_redist3d_f32(a.sweeps, n, n, n, sg8.ptr, f32.ptr, C.c_float(0.1))))
    out["reinit_f32"] = {"us_per_sweep": t * 1e3 / a.sweeps, "bytes_per_voxel": 9,
                         "GBps": 9.0 * G / (t * 1e-3 / a.sweeps) / 1e9}
    # connectivity: AD run of a.ad_iters iterations (iter = 2 .. term), solid phase
    rd = DeviceArray.from_host(np.asfortranarray(sdf, dtype=np.float64).ravel(order="F"))
    cn = DeviceArray(N, np.float64)
    t = timed(lambda: check(L.elsim_dev_accdiff_phase(n, n, n, rd.ptr, 0, a.ad_iters, cn.ptr)))
    it = a.ad_iters - 1
    # tiles whose inputs did not change since the previous iteration are
    # skipped, so this is an average over the run: the rate counts the 24 B
    # of every voxel in every iteration
    out["accdiff_f64"] = {"iterations": it, "ms": t, "us_per_iteration_avg": t * 1e3 / it, "bytes_per_voxel": 24,
                          "effective_GBps": 24.0 * N / (t * 1e-3 / it) / 1e9}
    lab, con = DeviceArray(N, np.int32), DeviceArray(N, np.int32)
    for ph in (0, 1):
        t = timed(lambda: check(L.elsim_dev_label_connected(n, n, n, rd.ptr, ph, lab.ptr, con.ptr)), reps=3)
        out[f"labels_phase{ph}"] = {"ms": t, "GBps_of_one_pass": 5.0 * N / (t * 1e-3) / 1e9}
    print(json.dumps(out), flush=True)
    if a.report:
        with open(a.report, "w") as fo:
            json.dump(out, fo, indent=1)
    for b in (sg64, sg8, f64, f32, rd, cn, lab, con):
        b.free()


if __name__ == "__main__":
    main()
