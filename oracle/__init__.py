"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so) and of the
reference's own level-set subroutines (oracle/_ref/libref_{ls,rm}.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
LIB_OMP = os.path.join(HERE, "liboracle_omp.so")  # bench.py's CPU baseline only
REF_LS = os.path.join(HERE, "_ref", "libref_ls.so")
REF_RM = os.path.join(HERE, "_ref", "libref_rm.so")

_o = None
_omp = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _o
    if _o is None:
        if not os.path.exists(LIB):
            build()
        _o = _setup(C.CDLL(LIB))
    return _o


def _setup(L):
    L.or_sbm_create.restype = C.c_void_p
    L.or_table.restype = C.c_double
    L.or_table.argtypes = [C.c_double, C.c_void_p, C.c_void_p, C.c_int]
    for f in ("or_sbm_step", "or_sbm_get", "or_sbm_set", "or_ch_update", "or_bv", "or_sbm_destroy",
              "or_sbm_scalars", "or_sbm_set_params"):
        getattr(L, f).argtypes = None
    return L


def omp_lib():
    """The same oracle built with OpenMP (loops on every host core): used only
    to time the CPU baseline, never as the checker."""
    global _omp
    if _omp is None:
        if not os.path.exists(LIB_OMP):
            build()
        _omp = _setup(C.CDLL(LIB_OMP))
        _omp.or_pcg_time.restype = C.c_double
        _omp.or_threads.restype = C.c_int
    return _omp


def pcg_time(op, kap, pse, rhs, dt, dx, niter):
    """Seconds for niter FP64 Jacobi-PCG iterations on the OpenMP build."""
    n1, n2, n3 = rhs.shape
    return omp_lib().or_pcg_time(op, n1, n2, n3, p(flat(kap)), p(flat(pse)), p(flat(rhs)), C.c_double(dt),
                                 C.c_double(dx), niter)


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def flat(a, dtype=np.float64):
    """Fortran-order flat copy (never a view: callees write into it)."""
    return np.array(np.asarray(a, dtype=dtype).ravel(order="F"), dtype=dtype, copy=True)


def unflat(a, shape):
    return a.reshape(shape, order="F")


# ------------------------------------------------------------ level set --
def nogrdbc3d(f):
    n1, n2, n3 = (s - 2 for s in f.shape)
    x = flat(f)
    lib().or_nogrdbc3d(n1, n2, n3, p(x))
    return unflat(x, f.shape)


def absgrd3duw(sg, f):
    n1, n2, n3 = (s - 2 for s in f.shape)
    g = np.zeros(n1 * n2 * n3)
    lib().or_absgrd3duw(n1, n2, n3, p(flat(sg)), p(flat(f)), p(g))
    return unflat(g, (n1, n2, n3))


def redist3d(lit, sg, f, dt=0.1):
    n1, n2, n3 = (s - 2 for s in f.shape)
    x = flat(f)
    lib().or_redist3d(lit, n1, n2, n3, p(flat(sg)), p(x), C.c_double(dt))
    return unflat(x, f.shape)


def redist3d_f32(lit, sg8, f, dt=np.float32(0.1)):
    n1, n2, n3 = (s - 2 for s in f.shape)
    x = flat(f, np.float32)
    lib().or_redist3d_f32(lit, n1, n2, n3, p(flat(sg8, np.int8)), p(x), C.c_float(dt))
    return unflat(x, f.shape)


def ls_frontend(tmpa, grid, sep, tr, n_ac=50):
    """LevelSet_3D's geometry front-end (LS:84-144) on the voxel image
    tmpa(yo, xo, zo): returns the ghosted sign field and the smoothed phase
    field on grid = (py, px, pz)."""
    yo, xo, zo = tmpa.shape
    py, px, pz = grid
    g = (py + 2) * (px + 2) * (pz + 2)
    sg, ds = np.zeros(g), np.zeros(g)
    rc = lib().or_ls_frontend(yo, xo, zo, p(flat(tmpa)), py, px, pz, sep, C.c_double(tr), n_ac, p(sg), p(ds))
    if rc != 0:
        raise ValueError("front-end geometry does not fit the voxel image")
    gs = (py + 2, px + 2, pz + 2)
    return unflat(sg, gs), unflat(ds, gs)


def time_control(xrf, scrnt, trgi, vsr0):
    """The SBM program's dt and Vsr for the next step (SBM:358-383)."""
    dt, vsr = C.c_double(), C.c_double()
    lib().or_time_control(C.c_double(xrf), C.c_double(scrnt), C.c_double(trgi), C.c_double(vsr0), C.byref(dt),
                          C.byref(vsr))
    return dt.value, vsr.value


def rm_phase(rdin, cn, phase):
    """RM_UCNCT_3D's sign block of one phase (RM:64-96, 115-136)."""
    n1, n2, n3 = rdin.shape
    sg = np.zeros((n1 + 2) * (n2 + 2) * (n3 + 2))
    lib().or_rm_phase(n1, n2, n3, p(flat(rdin)), p(flat(cn)), phase, p(sg))
    return unflat(sg, (n1 + 2, n2 + 2, n3 + 2))


def refine2(sdf):
    """The 2x refinement to FP32 between RM_UCNCT_3D and the SBM program."""
    n1, n2, n3 = sdf.shape
    out = np.zeros(8 * n1 * n2 * n3, np.float32)
    lib().or_refine2(n1, n2, n3, p(flat(sdf)), p(out))
    return unflat(out, (2 * n1, 2 * n2, 2 * n3))


def _ref(path):
    if not os.path.exists(path):
        return None
    return C.CDLL(path)


def ref_available():
    return os.path.exists(REF_LS) and os.path.exists(REF_RM)


def ref_redist3d(lit, sg, f, which="ls"):
    """The reference's ReDist3D_XYZ (LevelSet_3D.f90 or RM_UCNCT_3D.f90)."""
    L = _ref(REF_LS if which == "ls" else REF_RM)
    n1, n2, n3 = (s - 2 for s in f.shape)
    x = flat(f)
    ii = [C.c_int(v) for v in (lit, n1, n2, n3)]
    L.redist3d_xyz_(*[C.byref(v) for v in ii], p(flat(sg)), p(x))
    return unflat(x, f.shape)


def ref_absgrd3duw(sg, f, which="ls"):
    L = _ref(REF_LS if which == "ls" else REF_RM)
    n1, n2, n3 = (s - 2 for s in f.shape)
    g = np.zeros(n1 * n2 * n3)
    ii = [C.c_int(v) for v in (n1, n2, n3)]
    L.absgrd3duw_xyz_(*[C.byref(v) for v in ii], p(flat(sg)), p(flat(f)), p(g))
    return unflat(g, (n1, n2, n3))


def ref_nogrdbc3d(f, which="ls"):
    L = _ref(REF_LS if which == "ls" else REF_RM)
    n1, n2, n3 = (s - 2 for s in f.shape)
    x = flat(f)
    ii = [C.c_int(v) for v in (n1, n2, n3)]
    L.nogrdbc3d_(*[C.byref(v) for v in ii], p(x))
    return unflat(x, f.shape)


# --------------------------------------------------------- connectivity --
def accdiff_phase(rdin, phase, term=1501):
    n1, n2, n3 = rdin.shape
    cn = np.zeros(n1 * n2 * n3)
    lib().or_accdiff_phase(n1, n2, n3, p(flat(rdin)), phase, term, p(cn))
    return unflat(cn, rdin.shape)


def label_connected(rdin, phase):
    n1, n2, n3 = rdin.shape
    lab = np.zeros(n1 * n2 * n3, np.int32)
    con = np.zeros(n1 * n2 * n3, np.int32)
    lib().or_label_connected(n1, n2, n3, p(flat(rdin)), phase, p(lab), p(con))
    return unflat(lab, rdin.shape), unflat(con, rdin.shape)


# ------------------------------------------------------------------ SBM --
def table(c, cord, vals):
    return lib().or_table(float(c), p(cord), p(vals), len(cord))


def apply_op(op, kap, pse, x, dt, dx):
    n1, n2, n3 = (s - 2 for s in x.shape)
    out = np.zeros(n1 * n2 * n3)
    pse_f = flat(pse) if pse is not None else flat(np.zeros_like(kap))
    lib().or_apply_op(op, n1, n2, n3, p(flat(kap)), p(pse_f), p(flat(x)), C.c_double(dt), C.c_double(dx), p(out))
    return unflat(out, (n1, n2, n3))


def prj_bc(op, x):
    n1, n2, n3 = (s - 2 for s in x.shape)
    f = flat(x)
    lib().or_prj_bc(op, n1, n2, n3, p(f))
    return unflat(f, x.shape)


def ghost_mirror(x):
    n1, n2, n3 = (s - 2 for s in x.shape)
    f = flat(x)
    lib().or_ghost_mirror(n1, n2, n3, p(f))
    return unflat(f, x.shape)


def coef(params, which, cne, pse):
    """Cell coefficient of the step's operators from CnE and psi_E (ghosted):
    which 0 = CnE solve (SBM:474,478), 1 = RHE (SBM:579,583), 2 = phE solve
    (SBM:601,604).  Returns (coefficient, Dab)."""
    n1, n2, n3 = (s - 2 for s in cne.shape)
    out, dab = np.zeros(cne.size), np.zeros(cne.size)
    lib().or_coef(C.byref(params), n1, n2, n3, which, p(flat(cne)), p(flat(pse)), p(out), p(dab))
    return unflat(out, cne.shape), unflat(dab, cne.shape)


def rhe(pce, cne, dx):
    """RHE = dx^-2 S(pce, CnE) over the interior (SBM:591-596)."""
    n1, n2, n3 = (s - 2 for s in cne.shape)
    out = np.zeros(n1 * n2 * n3)
    lib().or_rhe(n1, n2, n3, p(flat(pce)), p(flat(cne)), C.c_double(dx), p(out))
    return unflat(out, (n1, n2, n3))


def avp(psp, pse, avp_min):
    """Interface density AvP (SBM:274-289) of ghosted psi_P, psi_E."""
    n1, n2, n3 = (s - 2 for s in psp.shape)
    out = np.zeros(n1 * n2 * n3)
    lib().or_avp(n1, n2, n3, p(flat(psp)), p(flat(pse)), C.c_double(avp_min), p(out))
    return unflat(out, (n1, n2, n3))


def solve_pot(kap, rhs, phi, bc_lo=1, bc_hi=1, dx=1.0, rtol=1e-8, maxit=100000):
    """The oracle's potential solve (the step's CG, SBM:657-714) with the
    ghost rules of elsim_solve_pot: mirror on the d1/d3 faces and a Neumann
    d2 face, stored values on a Dirichlet d2 plane.  Returns (phi, status,
    iterations, relres)."""
    op = 3 if (bc_lo and bc_hi) else (1 if bc_hi else 2)
    x = np.array(phi, dtype=np.float64)
    lo, hi = x[:, 0, :].copy(), x[:, -1, :].copy()
    x = ghost_mirror(x)
    if bc_lo:
        x[:, 0, :] = lo
    if bc_hi:
        x[:, -1, :] = hi
    n1, n2, n3 = rhs.shape
    xf = flat(x)
    it, rr = C.c_int(), C.c_double()
    st = lib().or_solve_system(op, n1, n2, n3, p(flat(kap)), p(flat(kap)), p(xf), p(flat(rhs)), C.c_double(1.0),
                               C.c_double(dx), 1, C.c_double(rtol), C.c_double(0.0), maxit, p(flat(kap)),
                               C.c_double(1.0), C.byref(it), C.byref(rr))
    x = unflat(xf, x.shape)
    lo, hi = x[:, 0, :].copy(), x[:, -1, :].copy()
    x = ghost_mirror(x)
    if bc_lo:
        x[:, 0, :] = lo
    if bc_hi:
        x[:, -1, :] = hi
    return x, st, it.value, rr.value


def psi_from_sdf(sdf, zeta):
    n1, n2, n3 = sdf.shape
    psi = np.zeros((n1 + 2) * (n2 + 2) * (n3 + 2))
    lib().or_psi_from_sdf(n1, n2, n3, p(flat(sdf, np.float32)), C.c_double(zeta), p(psi))
    return unflat(psi, (n1 + 2, n2 + 2, n3 + 2))


class OracleSbm:
    """The oracle's SBM state (same construction as elsim_sbm_create)."""

    GHOSTED = {0, 1, 3, 4, 5, 6, 8}

    def __init__(self, sdf_s, sdf_l, params, tab, L=None):
        self.L = L if L is not None else lib()
        self.shape = sdf_s.shape
        n1, n2, n3 = self.shape
        self.params = params
        self._s = flat(sdf_s, np.float32)
        self._l = flat(sdf_l, np.float32)
        self._t = np.ascontiguousarray(tab, dtype=np.float64)
        self.h = C.c_void_p(self.L.or_sbm_create(n1, n2, n3, C.byref(params), p(self._s), p(self._l),
                                                 self._t.size // 8, p(self._t)))

    def _shape(self, f):
        n1, n2, n3 = self.shape
        return (n1 + 2, n2 + 2, n3 + 2) if f in self.GHOSTED else (n1, n2, n3)

    def get(self, f):
        shp = self._shape(f)
        out = np.zeros(int(np.prod(shp)))
        self.L.or_sbm_get(self.h, f, p(out))
        return unflat(out, shp)

    def set(self, f, a):
        x = flat(a)
        self.L.or_sbm_set(self.h, f, p(x))

    def step(self):
        from elsim_amd import StepStats
        st = StepStats()
        self.L.or_sbm_step(self.h, C.byref(st))
        return st

    def ch_update(self):
        self.L.or_ch_update(self.h)

    def set_params(self, params):
        self.params = params
        self.L.or_sbm_set_params(self.h, C.byref(params))

    def scalars(self):
        v = [C.c_double() for _ in range(4)]
        it = C.c_int()
        self.L.or_sbm_scalars(self.h, *[C.byref(x) for x in v], C.byref(it))
        return dict(tpp=v[0].value, tpe=v[1].value, xrf=v[2].value, bve=v[3].value, iter=it.value)

    def close(self):
        if self.h:
            self.L.or_sbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
