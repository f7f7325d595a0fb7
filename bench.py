"""Benchmark: seconds per SBM electrochemical step at 256^3 on one MI355X.

One step = CH update -> CnE solve -> outer {Butler-Volmer -> phP solve -> phE
solve} -> current control (SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3.f90:397-840),
every CG to relative residual 1e-8.  The paths run on the same synthetic
microstructure from the same initial state:
  fast  - FP32 CG preconditioned by one multigrid V-cycle, FP64 reliable
          updates, extrapolated start plus the learned correction (the
          packaged network); also with the extrapolated and the plain start;
  jacobi - the same mixed solver with Jacobi preconditioning;
  fp64  - FP64 Jacobi-preconditioned CG, the reference's plain start;
  ref   - the reference CG (unpreconditioned FP64): cost per iteration only,
          since its solves do not reach 1e-8 within the reference's caps.
`value` is the fast path's seconds per step (lower is better).  Inputs are
resident in HBM before the timed region; W = 10 warm-up steps run first, then
K = 48 timed steps bracketed by device synchronisation.  The warm-up covers the
start-up transient (the first step starts from the flat initial state and needs
three or four coupling iterations).  The window is long because the cost of a
step is periodic: the current control changes the sign of its correction every
~14 steps at this size, the step after a change needs a second coupling
iteration and thousands of CG iterations, and the steps in between get cheaper
as the warm start's extrapolation improves (from ~1,300 down to ~50 phP
iterations).  Eight-step windows at different positions differed by 4x; 48
steps span three or four such cycles.

Also reported: the roofline of the dominant CG kernel, timed with HIP events
on the library stream inside the CG loop of one extra (untimed) step, and the
FP64 CPU oracle timed on a bounded sample on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
E_OP_PHP = 1

# HBM bytes per ghosted element moved by the FP32 CG kernels inside the CG
# loop at 256^3, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs over the
# launches that ran, in KiB per dispatch, scaled FETCH x 2.0 and WRITE x 1.0
# (the factors under which the pointwise kernels, whose bytes are known
# exactly, read their algorithmic bytes); profiles/cg_kernels_256_pmc.md.
# The flat kernels sweep the whole ghosted box, (n+2)^3 elements.
PMC_BYTES_PER_GHOSTED = {"upd32": (83892.2 * 2 + 67354.1) * 1024 / 258 ** 3,
                         "dir32": (117411.2 * 2 + 134168.1) * 1024 / 258 ** 3}
# the multigrid path's kernels, same method (profiles/mg_kernels_256_pmc.md):
# flat kernels over the ghosted box, the operator and the fine-level sweeps
# over the interior (bytes per interior voxel)
PMC_BYTES_PER_VOXEL_OP32 = (78082.8 * 2 + 67073.5) * 1024 / 256 ** 3
PMC_BYTES_PER_GHOSTED.update({"dir_mg": (100640.7 * 2 + 134168.1) * 1024 / 258 ** 3})
PMC_BYTES_PER_VOXEL_MG = {"mg_up": (191912.1 * 2 + 68654.6) * 1024 / 256 ** 3,
                          "mg_down": (231135.0 * 2 + 76764.8) * 1024 / 256 ** 3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_case(n, seed):
    import elsim_amd as E
    rng = np.random.default_rng(seed)
    target = float(rng.uniform(0.40, 0.60))
    sdf, ns, frac = E.sphere_pack(n, seed, 6.0, 16.0, target=target, hardcore=0.85, connect=True)
    return sdf, ns, frac, target


def matvec_roofline(n, reps=20):
    """The FP32 potential operator alone at n^3, each launch timed with HIP
    events on the library stream after a 1 GiB device copy has evicted its
    operands from the 256 MB last-level cache (as inside the CG loop, whose
    vectors do not fit there).  Algorithmic bytes: read p and kappa, write
    Ap = 12 B per voxel."""
    import elsim_amd as E
    L = E.lib()
    G = (n + 2) ** 3
    rng = np.random.default_rng(0)
    p = E.DeviceArray.from_host(rng.standard_normal(G).astype(np.float32))
    k = E.DeviceArray.from_host(rng.uniform(1e-7, 1.0, G).astype(np.float32))
    out = E.DeviceArray(n ** 3, np.float32)
    nb = 1 << 30
    fa, fb = E.DeviceArray(nb, np.uint8), E.DeviceArray(nb, np.uint8)
    E.check(L.elsim_memset(fa.ptr, 0, nb))
    for _ in range(3):
        E.check(L.elsim_dev_matvec_f32(E.OP_PHP, n, n, n, p.ptr, k.ptr, None, C.c_float(1.0), out.ptr, None))
    tot = 0.0
    ms = C.c_float()
    for _ in range(reps):
        E.check(L.elsim_memcpy_d2d(fb.ptr, fa.ptr, nb))
        E.check(L.elsim_timer_start())
        E.check(L.elsim_dev_matvec_f32(E.OP_PHP, n, n, n, p.ptr, k.ptr, None, C.c_float(1.0), out.ptr, None))
        E.check(L.elsim_timer_stop(C.byref(ms)))
        tot += ms.value
    t = tot * 1e-3 / reps
    return t, 12.0 * n ** 3 / t / 1e9


# Algorithmic bytes per interior voxel of the CG kernels (operands read and
# written once).  Operator: p, kappa in, Ap out (the CnE operator also reads
# psE: 4/8 B more).  FP64 update: x, p, r, w, dinv in; x, r out; FP64
# direction: r, p, dinv in; p out.  The mixed solver moves the solution
# update into the direction kernel: update r, w, dinv in, r out; direction y,
# r, p, dinv in, y, p out.  Its dinv is BF16 (2 B), the FP64 solver's FP64.
# The multigrid-preconditioned solver (n1 a multiple of 4) folds the residual
# update into the V-cycle's fine down-sweep: it reads r, w, dinv, kappa and
# writes the new r (to a second buffer) and the level-1 right-hand side
# (16 + 4 + 0.5 B per voxel); the up-sweep reads r, dinv, kappa and the
# level-1 iterate and writes z (12 + 0.5 + 4 B); the direction kernel reads
# y, p, z and writes y, p (20 B).  The coarse levels move about 1/7 of the
# fine bytes again (not counted in "vcycle").
KERNEL_BYTES = {"op32": 12, "upd32": 14, "dir32": 22, "op64": 24, "upd64": 56, "dir64": 32, "upd_mg": 20,
                "dir_mg": 20, "vcycle": 37, "mg_down": 20.5, "mg_up": 16.5}
KERNEL_NAMES = {"op32": "k_matvec4 (7-point operator, 4 voxels a lane, + p.Ap)", "upd32": "k_mx_update (r -= a w, r.z)",
                "dir32": "k_mx_dir (y += a p, p = z + b p)", "op64": "k_matvec<double>", "upd64": "k_pcg64_update",
                "dir64": "k_pcg64_dir", "upd_mg": "k_mg_update (r -= a w, x1 = omega dinv r)",
                "dir_mg": "k_mg_dir3 (y += a p, p = z + b p)",
                "vcycle": "V-cycle (fine-level bytes; all levels' kernels)",
                "mg_down": "k_mg_down4f (r = r_in - alpha w, x1 = omega dinv r formed in the loaders, "
                           "sigma (r - A x1) restricted in registers)",
                "mg_up": "k_mg_up4f (x2 = omega dinv r + oc P x_c in the loaders, z = x2 + omega dinv (r - A x2))"}


def profiled_step(E, s, mode, nsteps=3):
    """A few more steps with event brackets around the first CG iteration of
    every batch of the phP solves: mean in-loop duration of the operator,
    update and direction kernels (and the V-cycle and its fine-level
    sweeps)."""
    L = E.lib()
    P = E._lib
    E.check(L.elsim_prof_enable(2 + E_OP_PHP))
    prof = P.KernelProf()
    E.check(L.elsim_prof_read(C.byref(prof)))  # reset
    for _ in range(nsteps):
        s.step(mode)
    E.check(L.elsim_prof_read(C.byref(prof)))
    E.check(L.elsim_prof_enable(0))
    keys = ["op32", "upd32", "dir32", "op64", "upd64", "dir64", "vcycle", "mg_down", "mg_up"]
    if mode == E.MODE_MIXED_MG:
        keys[1], keys[2] = "upd_mg", "dir_mg"
    return {k: {"us": prof.mean_us(i), "launches": int(prof.count[i])} for i, k in enumerate(keys) if prof.count[i]}


def mg_settings(E):
    v = [C.c_double() for _ in range(3)]
    E.check(E.lib().elsim_mg_config_get(*[C.byref(x) for x in v]))
    return [x.value for x in v]


def nn_forward_time(E, net, n, reps=5):
    """Seconds per forward pass of the warm-start network at n^3 (HIP
    events on the library stream), on zero input, reading input channels
    0..7 as the SBM warm start does."""
    L = E.lib()
    x = E.DeviceArray((n ** 3 * 8,), np.uint16)
    E.check(L.elsim_memset(x.ptr, 0, x.nbytes))
    pred = E.DeviceArray((n ** 3,), np.float32)
    net.forward((n, n, n), x, pred, in_used=8)
    E.check(L.elsim_timer_start())
    for _ in range(reps):
        net.forward((n, n, n), x, pred, in_used=8)
    ms = C.c_float()
    E.check(L.elsim_timer_stop(C.byref(ms)))
    return ms.value * 1e-3 / reps


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sdf, prm, tab, cnp0, iters_per_step, target_s=12.0):
    """The FP64 oracle built with OpenMP (every host core) on a bounded sample
    of the same step: (1) one oracle step from the same state with every CG
    capped at one iteration and one coupling iteration -- all of the step's
    work except the CG iterations (CH, coefficients, RHE, Butler-Volmer,
    residual set-up); (2) FP64 Jacobi-PCG iterations of the phP system at full
    size (the algorithm of the FP64 GPU path), as many as fill ~target_s.
    s/step = (1) + (2) per iteration x the FP64 path's CG iterations per
    step."""
    import oracle as O
    from elsim_amd._lib import F_CNP, F_KAPP
    L = O.omp_lib()
    threads = int(L.or_threads())
    o = O.OracleSbm(sdf, -sdf, prm.copy(max_it_e=1, max_it_p=1, max_it_l=1, max_outer=1, skip_cne_first=0), tab,
                    L=L)
    o.set(F_CNP, cnp0)
    t0 = time.perf_counter()
    o.step()
    t_rest = time.perf_counter() - t0
    kap = o.get(F_KAPP)
    o.close()
    n1, n2, n3 = sdf.shape
    rhs = np.asfortranarray(np.random.default_rng(1).standard_normal((n1, n2, n3)) * 1e-3)
    t5 = O.pcg_time(E_OP_PHP, kap, kap, rhs, prm.dt, prm.dx, 5)
    k = int(min(2000, max(10, target_s / max(t5 / 5, 1e-6))))
    t = O.pcg_time(E_OP_PHP, kap, kap, rhs, prm.dt, prm.dx, k)
    per_it = t / k
    return {"value": t_rest + per_it * iters_per_step, "unit": "s/step", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"oracle (C, FP64, OpenMP on {threads} threads): one step with every CG capped at one "
                      f"iteration ({t_rest:.2f} s: CH, coefficients, RHE, Butler-Volmer, residual set-up) plus "
                      f"{k} FP64 Jacobi-PCG iterations of the phP system at {n1}^3 ({t:.2f} s, "
                      f"{per_it * 1e3:.1f} ms each) scaled by the FP64 GPU path's {iters_per_step:.0f} CG "
                      f"iterations per step"}


def reference_cg_cost(E, sdf, prm, tab, cnp0, cap=400):
    """Seconds per iteration of the reference CG (unpreconditioned FP64) on the
    same state after one step: its solves do not reach 1e-8 within the
    reference's caps at this size, so only the per-iteration cost is timed."""
    from elsim_amd._lib import F_CNP
    s = E.Sbm(sdf, -sdf, prm, tab)
    s.set(F_CNP, cnp0)
    s.step(E.MODE_MIXED)
    s.set_params(prm.copy(max_it_p=cap, max_it_l=cap, max_it_e=cap, max_outer=1))
    E.check(E.lib().elsim_synchronize())
    t0 = time.perf_counter()
    st = s.step(E.MODE_FP64)
    E.check(E.lib().elsim_synchronize())
    t = time.perf_counter() - t0
    s.close()
    its = st.it_cne + st.it_php + st.it_phe
    log(f"  reference CG: {its} iterations in {t:.2f} s, relres phP {st.relres_php:.2e} phE {st.relres_phe:.2e}")
    return {"s_per_iter": t / max(its, 1), "iters": its, "relres_php_after_cap": st.relres_php,
            "relres_phe_after_cap": st.relres_phe, "cap_per_solve": cap}


def run_path(E, sdf, prm, tab, mode, warmup, steps, cnp0, profile=True, net=None, fields=False):
    from elsim_amd._lib import F_CNE, F_CNP, F_PHE, F_PHP
    s = E.Sbm(sdf, -sdf, prm, tab)
    s.set(F_CNP, cnp0)
    if net is not None:
        s.set_warm_start(net)
    stats = []
    for w in range(warmup):
        st = s.step(mode)
        log(f"  warmup {w}: {st.as_dict()}")
    E.check(E.lib().elsim_synchronize())
    t0 = time.perf_counter()
    for k in range(steps):
        st = s.step(mode)
        stats.append(st.as_dict())
        log(f"  step {k}: {st.as_dict()}")
    E.check(E.lib().elsim_synchronize())
    t = time.perf_counter() - t0
    out_fields = {k: s.get(f)[1:-1, 1:-1, 1:-1] for k, f in (("cnp", F_CNP), ("cne", F_CNE), ("php", F_PHP),
                                                              ("phe", F_PHE))} if fields else None
    kern = profiled_step(E, s, mode) if profile else {}
    s.set_warm_start(None)
    if kern:
        log("  in-loop kernels: " + ", ".join(f"{k} {v['us']:.1f} us" for k, v in kern.items()))
    s.close()
    return t / steps, stats, kern, out_fields


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--paths", default="fast,fast_extrap,fast_plain,jacobi,jacobi_extrap,fp64,ref",
                    help="comma list of fast, fast_extrap, fast_plain (multigrid-preconditioned mixed solver), "
                         "jacobi, jacobi_extrap, jacobi_plain (Jacobi-preconditioned mixed solver), fp64, "
                         "fp64_same_start, ref")
    ap.add_argument("--weights", default=None, help="warm-start network (default: the packaged one)")
    ap.add_argument("--inner-rtol", type=float, default=None)
    ap.add_argument("--mg", default=None, help="omega,oc,delta of the multigrid path (default: the library's)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the operator timing and a 512 MiB device copy (for counter runs)")
    ap.add_argument("--cg-cap", type=int, default=0,
                    help="profiling runs only: cap every CG at this many iterations and the coupling loop at one "
                         "iteration (the printed line is then not the metric)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the path has no multi-GPU mode.",
              file=sys.stderr)
        sys.exit(2)
    import elsim_amd as E
    L = E.lib()
    E.check(L.elsim_init(0), "elsim_init")
    if a.mg:
        E.check(L.elsim_mg_config(*[float(v) for v in a.mg.split(",")]), "elsim_mg_config")
    n = a.n
    if a.roofline_only:
        t_mv, gbs = matvec_roofline(n, reps=10)
        nb = 512 << 20
        src, dst = E.DeviceArray(nb, np.uint8), E.DeviceArray(nb, np.uint8)
        E.check(L.elsim_memset(src.ptr, 1, nb))
        for _ in range(10):
            E.check(L.elsim_memcpy_d2d(dst.ptr, src.ptr, nb))
        E.check(L.elsim_synchronize())
        print(json.dumps({"matvec_us": t_mv * 1e6, "matvec_gbs": gbs, "copy_bytes": nb}))
        return
    t0 = time.perf_counter()
    sdf, ns, frac, target = make_case(n, a.seed)
    log(f"case: n={n} spheres={ns} solid={frac:.3f} (target {target:.3f}) gen {time.perf_counter() - t0:.1f}s")
    # the CnE solve resolves the step's change of the concentration to rtol
    # (its b_eff is dominated by psi_E c itself, so the b_eff rule would be
    # met at the start and CnE would do no work)
    prm = E.synthetic_params(cne_rule=1)
    tab = E.tables()
    cnp0 = E.initial_cnp(n, a.seed)

    if a.inner_rtol is not None:
        prm.inner_rtol = a.inner_rtol
    if a.cg_cap > 0:
        prm = prm.copy(max_it_e=a.cg_cap, max_it_p=a.cg_cap, max_it_l=a.cg_cap, max_outer=1)
    paths = a.paths.split(",")
    nan = float("nan")
    tfast, stf, t64, st64, ref_it, kf, k64, ffast, f64 = nan, [], nan, [], None, {}, {}, None, None
    from elsim_amd import nn as NN
    net = NN.WarmStartNet.load(a.weights or NN.DEFAULT_WEIGHTS)
    variants = {}
    # The multigrid paths extrapolate phi_P with the second-difference
    # Galerkin step but not the third: over three 256^3 packs x 48 steps the
    # third saved no iterations and cost ~0.45 ms a step (12.6 against 13.0
    # ms/step; linear extrapolation alone 15.4 ms/step).
    prm_mg = prm.copy(ws_order=2)
    if "fast" in paths:
        log("fast path (mixed precision, multigrid-preconditioned CG, learned warm start)")
        tfast, stf, kf, ffast = run_path(E, sdf, prm_mg.copy(ws_mode=E.WS_LEARNED), tab, E.MODE_MIXED_MG, a.warmup,
                                         a.steps, cnp0, net=net, fields=True)
        variants["learned"] = (tfast, stf)
    for key, wm, label in (("fast_extrap", E.WS_EXTRAP, "extrapolated"), ("fast_plain", E.WS_PLAIN, "plain")):
        if key in paths:
            log(f"mixed precision, multigrid, {label} warm start")
            t, st, _, _ = run_path(E, sdf, prm_mg.copy(ws_mode=wm), tab, E.MODE_MIXED_MG, a.warmup, a.steps, cnp0,
                                   profile=False)
            variants[label] = (t, st)
    jacobi = {}
    kj = {}
    for key, wm, label in (("jacobi", E.WS_LEARNED, "learned"), ("jacobi_extrap", E.WS_EXTRAP, "extrapolated"),
                           ("jacobi_plain", E.WS_PLAIN, "plain")):
        if key in paths:
            log(f"mixed precision, Jacobi-preconditioned CG, {label} warm start")
            t, st, k, _ = run_path(E, sdf, prm.copy(ws_mode=wm), tab, E.MODE_MIXED, a.warmup, a.steps, cnp0,
                                   profile=key == "jacobi", net=net if wm == E.WS_LEARNED else None)
            jacobi[label] = (t, st)
            kj.update(k)
    if "fp64" in paths:
        log("fp64 path (Jacobi-PCG, all FP64, plain warm start)")
        t64, st64, k64, f64 = run_path(E, sdf, prm.copy(ws_mode=E.WS_PLAIN), tab, E.MODE_FP64_PCG, a.warmup,
                                       a.steps, cnp0, fields=True)
    f64s = None
    if "fp64_same_start" in paths:
        log("fp64 path with the fast path's warm start (for the precision-only field error)")
        _, _, _, f64s = run_path(E, sdf, prm_mg.copy(ws_mode=E.WS_LEARNED), tab, E.MODE_FP64_PCG, a.warmup,
                                 a.steps, cnp0, profile=False, net=net, fields=True)
    t_nn = nn_forward_time(E, net, n) if n % 8 == 0 else nan
    log(f"warm-start network forward at {n}^3: {t_nn * 1e3:.2f} ms")
    if "ref" in paths:
        log("reference CG (unpreconditioned FP64): cost per iteration on capped solves")
        ref_it = reference_cg_cost(E, sdf, prm, tab, cnp0)

    t_mv, gbs = matvec_roofline(n)
    log(f"matvec f32 {n}^3 standalone, cold cache: {t_mv * 1e6:.1f} us, {gbs:.0f} GB/s")
    N = n ** 3
    kernels = {}
    for k, v in list(kf.items()) + list(k64.items()):
        kernels[k] = {"kernel": KERNEL_NAMES[k], "us": v["us"], "bytes_per_voxel": KERNEL_BYTES[k],
                      "GBps": KERNEL_BYTES[k] * N / (v["us"] * 1e-6) / 1e9, "launches_timed": v["launches"]}
    kernels_jacobi = {k: {"kernel": KERNEL_NAMES[k], "us": v["us"], "bytes_per_voxel": KERNEL_BYTES[k],
                          "GBps": KERNEL_BYTES[k] * N / (v["us"] * 1e-6) / 1e9, "launches_timed": v["launches"]}
                      for k, v in kj.items()}
    # dominant kernel of the fast path: largest share of the CG iteration.
    # The operator and the direction kernel take the same time to within a
    # few tenths of a microsecond, so a tie within 2 % goes to the operator,
    # the one further from its roofline (the conservative figure).
    single = {k: v for k, v in kf.items() if k != "vcycle"}
    dom = max(single, key=lambda k: single[k]["us"]) if single else None
    if dom and "op32" in kf and kf["op32"]["us"] >= 0.98 * kf[dom]["us"]:
        dom = "op32"
    if dom:
        if dom == "op32":
            traffic = PMC_BYTES_PER_VOXEL_OP32 * N
        elif dom in PMC_BYTES_PER_VOXEL_MG:
            traffic = PMC_BYTES_PER_VOXEL_MG[dom] * N
        else:
            traffic = PMC_BYTES_PER_GHOSTED[dom] * (n + 2) ** 3 if dom in PMC_BYTES_PER_GHOSTED else None
        roof = {"bound": "hbm", "achieved": kernels[dom]["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": kernels[dom]["GBps"] / HBM_PEAK_GBS, "traffic": traffic,
                "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE per launch, profiles/mg_kernels_256_pmc.md "
                                  "(operator and multigrid kernels) and profiles/cg_kernels_256_pmc.md",
                "kernel": kernels[dom]["kernel"],
                "kernel_us": kernels[dom]["us"], "bytes_per_launch": KERNEL_BYTES[dom] * N,
                "stencil": {"kernel": kernels["op32"]["kernel"], "us": kernels["op32"]["us"],
                            "GBps": kernels["op32"]["GBps"], "frac": kernels["op32"]["GBps"] / HBM_PEAK_GBS,
                            "traffic": PMC_BYTES_PER_VOXEL_OP32 * N, "bytes_per_launch": KERNEL_BYTES["op32"] * N,
                            "standalone_cold_us": t_mv * 1e6}}
    else:
        roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                "traffic": None, "kernel": "k_matvec4<phP> standalone, cold cache", "kernel_us": t_mv * 1e6}

    iters64 = float(np.mean([s["it_cne"] + s["it_php"] + s["it_phe"] for s in st64])) if st64 else nan
    itersf = float(np.mean([s["it_cne"] + s["it_php"] + s["it_phe"] for s in stf])) if stf else nan
    cpu = None
    if not a.no_cpu and st64:
        log("CPU baseline (OpenMP oracle, bounded sample)")
        cpu = cpu_baseline(sdf, prm, tab, cnp0, iters64)
        log(f"  {cpu['sample']}")
    # error of the fast path's fields against the FP64 path's after the same
    # steps from the same state (both solve every system to relres rtol)
    relerr = lambda f, g: {k: float(np.linalg.norm(f[k] - g[k]) / np.linalg.norm(g[k])) for k in f}  # noqa: E731
    field_err = relerr(ffast, f64) if ffast is not None and f64 is not None else None
    field_err_same = relerr(ffast, f64s) if ffast is not None and f64s is not None else None
    mean = lambda st, k: float(np.mean([x[k] for x in st])) if st else nan
    out = {
        "metric": "s/electrochem step @256³: FP64 CG vs FP32-refined+NN warm start, rel res 1e-8",
        "value": tfast,
        "unit": "s/step",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": tfast * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"sbm_electrochem_step_{n}^3 (nondimensional synthetic parameters)",
                   "grid": [n, n, n], "spheres": ns, "solid_fraction": round(frac, 4), "sphere_radius": [6, 16],
                   "hardcore": 0.85, "connect": True, "dt": prm.dt, "dx": prm.dx, "zeta": prm.zeta,
                   "F_over_RT": prm.cst1, "i0_table": 1.0, "i0_scale": prm.i0_scale, "alpha": 0.5,
                   "kappa": "psi_P", "D": "exp(-0.2 c)", "rtol": prm.rtol,
                   "cne_rule": "relative to the step's change (||r0||)", "cne_iters_per_step": mean(stf, "it_cne"),
                   "coupling_iters_per_step": {"fast": mean(stf, "n_outer"), "fp64": mean(st64, "n_outer")},
                   "fast_path": "fp32 cg preconditioned by a multigrid v-cycle + fp64 reliable updates + learned "
                                "warm start (CnE: fp32 jacobi-pcg)",
                   "multigrid": dict(zip(("omega", "oc", "delta"), mg_settings(E))),
                   "ws_order_fast": 2,
                   "learned_warm_start": True},
        "field_rel_l2_vs_fp64": field_err,
        "field_rel_l2_vs_fp64_same_start": field_err_same,
        "field_error_note": "vs_fp64: against the FP64 path with the reference's plain start, whose coupling loop "
                            "exits on CvgL with CvgP up to ~1e-6 (SBM:627); same_start: against FP64 Jacobi-PCG "
                            "with the fast path's warm start, i.e. the error of the reduced precision alone",
        "fp64_s_per_step": t64,
        "fp64_path": "fp64 jacobi-pcg (same preconditioner, all FP64), plain warm start as the reference",
        "speedup_vs_fp64": t64 / tfast,
        "reference_cg": ref_it,
        "warm_start": {
            k: {"s_per_step": v[0],
                "php_iters_per_step": float(np.mean([x["it_php"] for x in v[1]])) if v[1] else nan,
                "cg_iters_per_step": float(np.mean([x["it_cne"] + x["it_php"] + x["it_phe"] for x in v[1]]))
                if v[1] else nan}
            for k, v in variants.items()},
        "jacobi_preconditioned": {
            k: {"s_per_step": v[0],
                "php_iters_per_step": float(np.mean([x["it_php"] for x in v[1]])) if v[1] else nan,
                "cg_iters_per_step": float(np.mean([x["it_cne"] + x["it_php"] + x["it_phe"] for x in v[1]]))
                if v[1] else nan}
            for k, v in jacobi.items()},
        "nn_forward_ms": t_nn * 1e3,
        "nn_weights": os.path.relpath(a.weights or NN.DEFAULT_WEIGHTS, ROOT),
        "cg_iters_per_step": {"fp64": iters64, "fast": itersf},
        "voxel_updates_per_s": {"fp64": N * iters64 / t64, "fast": N * itersf / tfast},
        "roofline": roof,
        "kernels_in_loop": kernels,
        "kernels_in_loop_jacobi": kernels_jacobi,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
