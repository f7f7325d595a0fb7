"""Scale-up run at 512^3: ellipsoid packing (semi-axes 8-24 voxels, major
axis within 15 degrees of the through-thickness axis d2) with a random-walk
tunnel, connectivity clean-up by component labelling (the reference's
RM_UCNCT_3D step: solid not connected to the current collector and
electrolyte not connected to the separator are removed), then coupled
electrochemical steps with the mixed-precision solver (multigrid-preconditioned,
or Jacobi-preconditioned with --solver jacobi) and the learned warm start.
Reports per-step iterations and times, the device memory footprint and the
in-loop HBM rates of the CG kernels.

    python -m elsim_amd.scaleup --n 512 --steps 20 --report scaleup.json
"""
import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

import elsim_amd as E
from elsim_amd import nn as NN
from elsim_amd._lib import F_CNP, KernelProf

# algorithmic bytes per voxel: the Jacobi-preconditioned solver keeps its
# inverse in BF16; the multigrid path's update and direction move 20 B, its
# fine-level sweeps 12.5 and 20.5 B (with the level-1 vectors)
KERNEL_BYTES = {"op32": 12, "upd32": 14, "dir32": 22, "upd_mg": 20, "dir_mg": 20, "mg_down": 12.5, "mg_up": 20.5}


def log(*a):
    print(*a, flush=True)


def mem_used():
    free, total = C.c_size_t(), C.c_size_t()
    E.check(E.lib().elsim_mem_info(C.byref(free), C.byref(total)))
    return total.value - free.value, total.value


def clean(sdf, dcut):
    """Drop solid not face-connected (phase 0) and fill electrolyte not
    face-connected (phase 1), from the GPU component labelling."""
    n1, n2, n3 = sdf.shape
    L = E.lib()
    rd = E.DeviceArray.from_host(sdf.astype(np.float64).ravel(order="F"))
    lab = E.DeviceArray((sdf.size,), np.int32)
    con = E.DeviceArray((sdf.size,), np.int32)
    out = sdf.copy()
    stats = {}
    for phase in (0, 1):
        t0 = time.perf_counter()
        E.check(L.elsim_dev_label_connected(n1, n2, n3, rd.ptr, phase, lab.ptr, con.ptr))
        E.check(L.elsim_synchronize())
        dt = time.perf_counter() - t0
        lb = lab.to_host().reshape(sdf.shape, order="F")
        cn = con.to_host().reshape(sdf.shape, order="F").astype(bool)
        inphase = lb >= 0
        drop = inphase & ~cn
        if phase == 0:
            out[drop] = -dcut
        else:
            out[drop] = 0.5
        stats[f"phase{phase}"] = {"components": int(len(np.unique(lb[inphase]))), "dropped_voxels": int(drop.sum()),
                                  "label_s": dt}
    return out, stats


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--ws", choices=["plain", "extrap", "learned"], default="learned")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--solver", choices=["mg", "jacobi"], default="mg")
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    E.check(E.lib().elsim_init(0), "elsim_init")
    n, dcut = a.n, 12.0
    base, total = mem_used()
    t0 = time.perf_counter()
    sdf, ns, frac = E.ellipsoid_pack(n, a.seed, 8.0, 24.0, cone_deg=15.0, target=0.5, overlap=0.3, dcut=dcut)
    sdf = E.carve_tunnel(sdf, a.seed, radius=3.0)
    t_gen = time.perf_counter() - t0
    sdf, cstats = clean(sdf, dcut)
    log(f"{n}^3: {ns} ellipsoids, solid {frac:.3f}, generated in {t_gen:.1f} s; clean-up {cstats}")
    wsm = {"plain": E.WS_PLAIN, "extrap": E.WS_EXTRAP, "learned": E.WS_LEARNED}[a.ws]
    # as the 256^3 benchmark: the CnE solve resolves the step's change to rtol
    mg = a.solver == "mg"
    mode = E.MODE_MIXED_MG if mg else E.MODE_MIXED
    # the multigrid path extrapolates with the second-difference Galerkin step only (as the benchmark)
    s = E.Sbm(sdf, -sdf, E.synthetic_params(ws_mode=wsm, cne_rule=1, ws_order=2 if mg else 0), E.tables())
    s.set(F_CNP, E.initial_cnp(n, a.seed))
    net = None
    if a.ws == "learned":
        net = NN.WarmStartNet.load(a.weights or NN.DEFAULT_WEIGHTS)
        s.set_warm_start(net)
    steps = []
    for k in range(a.steps):
        t1 = time.perf_counter()
        st = s.step(mode)
        E.check(E.lib().elsim_synchronize())
        d = st.as_dict()
        d["wall_s"] = time.perf_counter() - t1
        steps.append(d)
        log(f"  step {k}: phP {st.it_php} phE {st.it_phe} CnE {st.it_cne} outer {st.n_outer} "
            f"relres {st.relres_php:.1e}/{st.relres_phe:.1e} ws {st.ws_used} {d['wall_s']:.2f} s")
    used, _ = mem_used()
    # in-loop kernel rates on one more step
    L = E.lib()
    prof = KernelProf()
    E.check(L.elsim_prof_enable(1))
    E.check(L.elsim_prof_read(C.byref(prof)))
    for _ in range(2):
        s.step(mode)
    E.check(L.elsim_prof_read(C.byref(prof)))
    E.check(L.elsim_prof_enable(0))
    N = n ** 3
    kern = {}
    slots = (("op32", 0), ("upd_mg" if mg else "upd32", 1), ("dir_mg" if mg else "dir32", 2), ("mg_down", 7),
             ("mg_up", 8))
    for k, i in slots:
        if prof.count[i]:
            us = prof.mean_us(i)
            kern[k] = {"us": us, "GBps": KERNEL_BYTES[k] * N / (us * 1e-6) / 1e9, "launches": int(prof.count[i])}
    log(f"  in-loop kernels: {kern}")
    rep = {"config": {"grid": [n, n, n], "shapes": ns, "solid_fraction": frac, "semi_axes": [8, 24], "cone_deg": 15,
                      "tunnel_radius": 3, "seed": a.seed, "warm_start": a.ws,
                      "mode": ("fp32 multigrid-pcg" if mg else "fp32 jacobi-pcg") + " + fp64 reliable updates",
                      "cne_rule": "relative to the step's change (||r0||)"},
           "generate_s": t_gen, "cleanup": cstats, "steps": steps,
           "s_per_step_after_first": float(np.mean([x["wall_s"] for x in steps[1:]])) if len(steps) > 1 else None,
           "s_per_step_last_half": float(np.mean([x["wall_s"] for x in steps[len(steps) // 2:]])) if steps else None,
           "device_memory_bytes": used - base, "device_memory_total": total, "kernels_in_loop": kern}
    log(json.dumps({k: v for k, v in rep.items() if k != "steps"}))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    s.close()
    if net is not None:
        net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
