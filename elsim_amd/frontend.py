"""Geometry front-end of LevelSet_3D (LevelSet_3D.f90:84-153) on the GPU.

`ls_frontend` is one frame of the reference's frame loop up to the
reinitialisation (elsim_ls_frontend: reorient the voxel image, carve the two
tunnels, stack the thickness copies, Allen-Cahn smoothing, sign function);
`ls_frame` adds the reinitialisation ReDist3D_XYZ(175) that turns it into the
signed distance function the frame writes (LS:147-151).

Run as a module it measures one frame at the reference's sizes (voxel image
500 x 520 x 170, grid 93 x (50 + 4 * 170) x 160, 50 smoothing iterations,
175 reinitialisation sweeps) on a synthetic voxel image, with the arrays
resident on the device, and prints one JSON line:

    python -m elsim_amd.frontend [--tr 19.171] [--report out.json]
"""
import argparse
import ctypes as C
import json
import time

import numpy as np

from ._lib import DeviceArray, check, lib, ptr
from .sbm import sphere_pack

# the reference's sizes (LS:9-11, 110-121, 126, 148)
REF_IMAGE = (500, 520, 170)
REF_GRID = (93, 50 + 4 * 170, 160)
REF_SEP = 50
REF_NAC = 50
REF_LIT = 175
# TuRd(3..13) as the reference stores them: REAL(4) literals in a REAL(8)
# DATA statement (LS:54-57)
TURD = [float(np.float32(v)) for v in (0, 6.3826e+00, 1.2764e+01, 1.9171e+01, 2.5566e+01, 3.1941e+01, 3.8331e+01,
                                        4.4729e+01, 5.1114e+01, 5.7506e+01, 6.3890e+01, 7.0280e+01, 7.6674e+01,
                                        8.3061e+01)]


def _ip(v):
    return C.byref(C.c_int(int(v)))


def ls_frontend(tmpa, grid, sep, tr, n_ac=REF_NAC):
    """Host arrays in and out: tmpa(yo, xo, zo) FP64 -> (sgF, dsF), ghosted
    (py+2, px+2, pz+2), Fortran order."""
    tmpa = np.asarray(tmpa, dtype=np.float64)
    yo, xo, zo = tmpa.shape
    py, px, pz = grid
    gs = (py + 2, px + 2, pz + 2)
    sg, ds = np.zeros(gs, order="F"), np.zeros(gs, order="F")
    ta = np.asfortranarray(tmpa)
    check(lib().elsim_ls_frontend(_ip(yo), _ip(xo), _ip(zo), ptr(ta), _ip(py), _ip(px), _ip(pz), _ip(sep),
                                  C.byref(C.c_double(tr)), _ip(n_ac), ptr(sg), ptr(ds)), "ls_frontend")
    return sg, ds


def ls_frame(tmpa, grid, sep, tr, n_ac=REF_NAC, lit=REF_LIT):
    """The frame's signed distance function: front-end, dsF = 0, then
    ReDist3D_XYZ(lit) with the LS time step 0.1d0 (LS:146-148)."""
    sg, _ = ls_frontend(tmpa, grid, sep, tr, n_ac)
    f = np.zeros_like(sg)
    py, px, pz = grid
    check(lib().elsim_redist3d_xyz(_ip(lit), _ip(py), _ip(px), _ip(pz), ptr(sg), ptr(f), C.byref(C.c_double(0.1))),
          "redist3d")
    return sg, f


def synthetic_voxels(shape, seed=1):
    """A 0/1 voxel image of the reference's layout (yo, xo, zo): a seeded
    sphere pack, 1 inside the particles (the reference's image file is not
    in its repository)."""
    sdf, _, _ = sphere_pack(shape, seed, 4.0, 12.0, target=0.5, hardcore=0.85, connect=False)
    return (sdf > 0).astype(np.float64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--tr", type=float, default=TURD[3])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--report", default=None)
    a = ap.parse_args()
    L = lib()
    check(L.elsim_init(0), "elsim_init")
    yo, xo, zo = REF_IMAGE
    py, px, pz = REF_GRID
    t0 = time.perf_counter()
    tmpa = synthetic_voxels(REF_IMAGE, a.seed)
    t_gen = time.perf_counter() - t0
    G = (py + 2) * (px + 2) * (pz + 2)
    N = py * px * pz
    dt = DeviceArray.from_host(np.asfortranarray(tmpa).ravel(order="F"))
    sg, ds, f = DeviceArray(G, np.float64), DeviceArray(G, np.float64), DeviceArray(G, np.float64)
    ms = C.c_float()
    res = {}
    for n_ac in (0, REF_NAC):  # the difference isolates the smoothing iterations
        best = None
        for _ in range(3):
            check(L.elsim_timer_start())
            check(L.elsim_dev_ls_frontend(yo, xo, zo, dt.ptr, py, px, pz, REF_SEP, C.c_double(a.tr), n_ac, sg.ptr,
                                          ds.ptr))
            check(L.elsim_timer_stop(C.byref(ms)))
            best = ms.value if best is None else min(best, ms.value)
        res[n_ac] = best
    check(L.elsim_memset(f.ptr, 0, f.nbytes))
    check(L.elsim_timer_start())
    check(L.elsim_dev_redist3d_f64(REF_LIT, py, px, pz, sg.ptr, f.ptr, C.c_double(0.1)))
    check(L.elsim_timer_stop(C.byref(ms)))
    t_re = ms.value
    sgh, dsh, fh = sg.to_host(), ds.to_host(), f.to_host()
    ac_ms = (res[REF_NAC] - res[0]) / REF_NAC
    ac_bytes = 16.0 * N  # read f, write the new f (interior); ghosts are 1.4 % more
    sweep_bytes = 24.0 * N  # read sg, f, write f
    out = {"workload": "levelset_frontend_frame", "image": list(REF_IMAGE), "grid": list(REF_GRID), "tr": a.tr,
           "n_ac": REF_NAC, "lit": REF_LIT, "generate_s": t_gen,
           "frontend_ms": res[REF_NAC], "build_and_sign_ms": res[0],
           "allen_cahn_ms_per_iter": ac_ms, "allen_cahn_GBps": ac_bytes / (ac_ms * 1e-3) / 1e9,
           "reinit_ms": t_re, "reinit_ms_per_sweep": t_re / REF_LIT,
           "reinit_GBps": sweep_bytes / (t_re / REF_LIT * 1e-3) / 1e9,
           "frame_ms": res[REF_NAC] + t_re,
           "sign_positive_fraction": float((sgh > 0).mean()),
           "sign_matches_field": bool(np.array_equal(sgh, np.where(dsh > 0.5, 1.0, -1.0))),
           "sdf_range": [float(fh.min()), float(fh.max())]}
    print(json.dumps(out), flush=True)
    if a.report:
        with open(a.report, "w") as fo:
            json.dump(out, fo, indent=1)
    for b in (dt, sg, ds, f):
        b.free()


if __name__ == "__main__":
    main()
