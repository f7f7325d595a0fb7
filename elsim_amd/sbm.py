"""Host-side mirror of the SBM solver core: synthetic problem set-up and a thin
object over the device-resident state.

Array convention (as the reference): Fortran order, d1 contiguous, ghosted
arrays of shape (n1+2, n2+2, n3+2); numpy arrays here are Fortran-ordered
views of the flat buffers.
"""
import ctypes as C

import numpy as np

from ._lib import (GHOSTED_FIELDS, MODE_FP64, STOP_RELRES, ElsimError, SbmParams, StepStats, check, iref, lib,
                   ptr)

T_MINUS = 4.0e-6 / (4.0e-6 + 1.25e-6)  # SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3.f90:121


def reference_params():
    """The reference's own constants (SBM:88-136), REAL(4) kinds resolved."""
    p = SbmParams()
    lib().elsim_sbm_default_params(C.byref(p))
    return p


def time_control(xrf, scrnt, trgi, vsr0):
    """dt and Vsr of the next step from the reference's time-loop control
    (SBM:358-383): Vsr0 scaled by five tiers of |trgI - sCrnt| / |trgI|, dt and
    a further factor by three tiers of Xrf."""
    dt, vsr = C.c_double(), C.c_double()
    d = C.c_double
    check(lib().elsim_sbm_time_control(C.byref(d(xrf)), C.byref(d(scrnt)), C.byref(d(trgi)), C.byref(d(vsr0)),
                                       C.byref(dt), C.byref(vsr)), "time_control")
    return dt.value, vsr.value


def synthetic_params(**kw):
    """Nondimensional parameters of the synthetic electrochemical step (h = 1
    voxel): psi = 0.5(1+tanh(d/2)), alpha = 0.5, F/RT = 1, dt = 1e-3, every CG to
    relative residual 1e-8.  The exchange current is the table value 1 times
    i0_scale = 5e-6 (the reference scales by 1e-3, SBM:617): it keeps the
    kinetic-to-conductive ratio a i0 (F/RT) L^2 / kappa near the reference's
    physical regime (~0.04) up to 512^3, where the block Gauss-Seidel coupling
    loop (SBM:627) contracts; at i0 = 1e-3 it diverges beyond ~64^3."""
    tm = T_MINUS
    p = SbmParams(
        dx=1.0, dt=1.0e-3, zeta=2.0, psp_off=1.0e-7, pse_off=1.0e-6, avp_min=1.0e-2, eps=1.0, rho=1.0,
        ch_psi_min=1.0e-5, frd=1.0, cst1=1.0, t_minus=tm, d0=1.0, tc1=(2 * tm - 1.0) / (2 * tm * (1.0 - tm)),
        tc2=1.0 / (2 * tm * (1.0 - tm)), dab_c0=0.0, dab_c1=0.2, dab_c2=0.0, dab_c3=0.0, kap_scale=1.0,
        i0_scale=5.0e-6, kfw_cref=1.0, bvp=0.0, bve=-0.5, vsr=0.1, trgi=0.0, cne_target=1.0, rtol=1.0e-8,
        crt_e=1.0e-7, crt_p=1.0e-7, crt_l=1.0e-6, outer_tol=1.0e-8, inner_rtol=0.5, mass_corr_period=100,
        skip_cne_first=1, stop_rule=STOP_RELRES, max_it_e=50000, max_it_p=30000, max_it_l=100000, max_outer=30)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def sphere_pack(n, seed, rmin, rmax, target=0.5, hardcore=0.85, connect=True, dcut=48.0):
    """Interior FP32 SDF (Fortran order) of a seeded sphere pack."""
    n1, n2, n3 = (n, n, n) if np.isscalar(n) else n
    sdf = np.zeros(n1 * n2 * n3, np.float32)
    ns, fr = C.c_int(), C.c_double()
    check(lib().elsim_gen_sphere_pack(n1, n2, n3, seed, rmin, rmax, target, hardcore, int(connect), dcut,
                                      ptr(sdf), C.byref(ns), C.byref(fr)), "gen_sphere_pack")
    return sdf.reshape((n1, n2, n3), order="F"), ns.value, fr.value


def ellipsoid_pack(n, seed, amin=8.0, amax=24.0, cone_deg=15.0, target=0.5, overlap=0.0, dcut=48.0):
    """Interior FP32 SDF (Fortran order) of seeded random ellipsoids aligned
    within cone_deg of d2; overlap = admitted solid fraction of a new shape."""
    n1, n2, n3 = (n, n, n) if np.isscalar(n) else n
    sdf = np.zeros(n1 * n2 * n3, np.float32)
    ns, fr = C.c_int(), C.c_double()
    check(lib().elsim_gen_ellipsoid_pack(n1, n2, n3, seed, amin, amax, cone_deg, target, overlap, dcut, ptr(sdf),
                                         C.byref(ns), C.byref(fr)), "gen_ellipsoid_pack")
    return sdf.reshape((n1, n2, n3), order="F"), ns.value, fr.value


def carve_tunnel(sdf, seed, radius=3.0):
    n1, n2, n3 = sdf.shape
    flat = np.asfortranarray(sdf).ravel(order="F").copy()
    check(lib().elsim_gen_tunnel(n1, n2, n3, seed, radius, ptr(flat)), "gen_tunnel")
    return flat.reshape((n1, n2, n3), order="F")


def potential_problem(n, seed, rmin=4.0, rmax=8.0, target=0.5, xi=2.0, sigma=1.0, i0=1.0, alpha=0.5, eta=0.1,
                      phase="solid"):
    """One potential solve on a sphere pack: kappa = psi sigma + 1e-7 (the
    reference's psi offset, SBM:247) with psi = 0.5(1+tanh(d/xi)) of the
    chosen phase, rhs = |grad psi| j with the Butler-Volmer current
    j = i0 (exp(alpha eta) - exp(-(1-alpha) eta)) at the fixed overpotential eta
    (F/RT = 1), phi = 0 / 1 in the low / high d2 ghost planes (Dirichlet) and 0.5
    inside.  The spheres do not overlap (hard core 1).  Returns (kappa, rhs,
    phi0, info): kappa, phi0 ghosted, rhs interior, Fortran order."""
    sdf, ns, frac = sphere_pack(n, seed, rmin, rmax, target=target, hardcore=1.0, connect=False)
    d = sdf.astype(np.float64) if phase == "solid" else -sdf.astype(np.float64)
    psi = np.pad(0.5 * (1.0 + np.tanh(d / xi)), 1, mode="edge")
    kap = psi * sigma + 1.0e-7
    g = [(np.roll(psi, -1, a) - np.roll(psi, 1, a))[1:-1, 1:-1, 1:-1] / 2 for a in range(3)]
    grad = np.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2)
    j = i0 * (np.exp(alpha * eta) - np.exp(-(1.0 - alpha) * eta))
    rhs = np.asfortranarray(grad * j)
    phi = np.full(psi.shape, 0.5)
    phi[:, 0, :] = 0.0
    phi[:, -1, :] = 1.0
    return np.asfortranarray(kap), rhs, np.asfortranarray(phi), {"spheres": ns, "solid_fraction": frac}


def solve_pot(kap, rhs, phi, bc_lo=1, bc_hi=1, mode=MODE_FP64, dx=1.0, rtol=1e-8, maxit=100000):
    """elsim_solve_pot on host arrays: returns (phi, status, iterations,
    relres); phi's Dirichlet ghost planes hold the boundary values."""
    n1, n2, n3 = rhs.shape
    k = np.ascontiguousarray(np.asarray(kap, dtype=np.float64).ravel(order="F"))
    r = np.ascontiguousarray(np.asarray(rhs, dtype=np.float64).ravel(order="F"))
    x = np.array(np.asarray(phi, dtype=np.float64).ravel(order="F"), copy=True)  # the solve writes into it
    it, rr = C.c_int(), C.c_double()
    rc = lib().elsim_solve_pot(iref(n1), iref(n2), iref(n3), C.byref(C.c_double(dx)), ptr(k), ptr(r), ptr(x),
                               iref(bc_lo), iref(bc_hi), iref(mode), C.byref(C.c_double(rtol)), iref(maxit),
                               C.byref(it), C.byref(rr))
    check(rc, "elsim_solve_pot")
    return x.reshape(phi.shape, order="F"), rc, it.value, rr.value


def tables(sz=101):
    t = np.zeros(8 * sz)
    check(lib().elsim_gen_tables(sz, ptr(t)), "gen_tables")
    return t


def initial_cnp(n, seed, mean=0.5, amp=0.01):
    """CnP over the ghosted box: mean +- amp uniform noise, mirrored ghosts."""
    n1, n2, n3 = (n, n, n) if np.isscalar(n) else n
    rng = np.random.default_rng(seed)
    c = mean + rng.uniform(-amp, amp, (n1, n2, n3))
    return np.pad(c, 1, mode="edge")


class Sbm:
    """Device-resident SBM state (elsim_sbm_* entry points)."""

    def __init__(self, sdf_s, sdf_l, params, tab=None):
        self.shape = sdf_s.shape
        n1, n2, n3 = self.shape
        self.params = params
        tab = tables() if tab is None else np.ascontiguousarray(tab, dtype=np.float64)
        self.sz = tab.size // 8
        s = np.ascontiguousarray(np.asfortranarray(sdf_s, dtype=np.float32).ravel(order="F"))
        l = np.ascontiguousarray(np.asfortranarray(sdf_l, dtype=np.float32).ravel(order="F"))
        self.h = C.c_int64(0)
        check(lib().elsim_sbm_create(iref(n1), iref(n2), iref(n3), C.byref(params), ptr(s), ptr(l),
                                     iref(self.sz), ptr(tab), C.byref(self.h)), "elsim_sbm_create")

    def _size(self, field):
        n1, n2, n3 = self.shape
        return ((n1 + 2, n2 + 2, n3 + 2) if field in GHOSTED_FIELDS else (n1, n2, n3))

    def get(self, field):
        shp = self._size(field)
        out = np.empty(int(np.prod(shp)))
        check(lib().elsim_sbm_get(C.byref(self.h), iref(field), ptr(out)), "elsim_sbm_get")
        return out.reshape(shp, order="F")

    def set(self, field, a):
        shp = self._size(field)
        a = np.asarray(a, dtype=np.float64)
        assert a.shape == shp, (a.shape, shp)
        flat = np.ascontiguousarray(a.ravel(order="F"))
        check(lib().elsim_sbm_set(C.byref(self.h), iref(field), ptr(flat)), "elsim_sbm_set")

    def field_ptr(self, field):
        p = C.c_void_p()
        check(lib().elsim_sbm_field_ptr(C.byref(self.h), iref(field), C.byref(p)), "field_ptr")
        return p

    def scalars(self):
        v = [C.c_double() for _ in range(4)]
        it = C.c_int()
        check(lib().elsim_sbm_scalars(C.byref(self.h), *[C.byref(x) for x in v], C.byref(it)), "scalars")
        return dict(tpp=v[0].value, tpe=v[1].value, xrf=v[2].value, bve=v[3].value, iter=it.value)

    def set_params(self, params):
        self.params = params
        check(lib().elsim_sbm_set_params(C.byref(self.h), C.byref(params)), "set_params")

    def step(self, mode=MODE_FP64):
        st = StepStats()
        rc = lib().elsim_sbm_step(C.byref(self.h), iref(mode), C.byref(st))
        check(rc, "elsim_sbm_step")
        return st

    def set_warm_start(self, net):
        """Attach a WarmStartNet (None: detach); the state keeps a reference."""
        self._ws = net
        nh = C.byref(net.h) if net is not None else None
        check(lib().elsim_sbm_set_warm_start(C.byref(self.h), nh), "elsim_sbm_set_warm_start")

    def ws_record(self, x_dev, start_dev):
        """Arm recording of the next step's warm-start network input (device
        bf16 [N][16]) and extrapolated start (device FP64 [N]); None disarms."""
        check(lib().elsim_sbm_ws_record(C.byref(self.h), x_dev.ptr if x_dev is not None else None,
                                        start_dev.ptr if start_dev is not None else None), "elsim_sbm_ws_record")

    def close(self):
        if self.h is not None and self.h.value:
            lib().elsim_sbm_destroy(C.byref(self.h))
        self.h = None
        self._ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
