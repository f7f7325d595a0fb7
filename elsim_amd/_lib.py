"""ctypes binding of libelsim_hip.so.

The library is loaded from elsim_amd/lib/ (built in-tree by elsim_amd.build).
There is no fallback: if the library is missing or no GPU is present, calls
raise ElsimError.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libelsim_hip.so")

OK, NOT_CONVERGED = 0, 1
ERR_ARG, ERR_HIP, ERR_NAN, ERR_ALLOC = -1, -2, -3, -4
STOP_REFERENCE, STOP_RELRES = 0, 1
MODE_FP64, MODE_MIXED, MODE_FP64_PCG, MODE_MIXED_MG = 0, 1, 2, 3
OP_CNE, OP_PHP, OP_PHE = 0, 1, 2
WS_PLAIN, WS_EXTRAP, WS_LEARNED = 0, 1, 2
F_PSP, F_PSE, F_AVP, F_CNP, F_CNE, F_PHP, F_PHE, F_RXN, F_KAPP = range(9)
GHOSTED_FIELDS = {F_PSP, F_PSE, F_CNP, F_CNE, F_PHP, F_PHE, F_KAPP}


class ElsimError(RuntimeError):
    pass


_PARAM_DOUBLES = ("dx dt zeta psp_off pse_off avp_min eps rho ch_psi_min frd cst1 t_minus d0 tc1 tc2 "
                  "dab_c0 dab_c1 dab_c2 dab_c3 kap_scale i0_scale kfw_cref bvp bve vsr trgi cne_target "
                  "rtol crt_e crt_p crt_l outer_tol inner_rtol").split()
_PARAM_INTS = ("mass_corr_period skip_cne_first stop_rule max_it_e max_it_p max_it_l max_outer ws_mode cne_rule "
               "ws_order").split()


class SbmParams(C.Structure):
    """Mirror of elsim_sbm_params (include/elsim.h)."""
    _fields_ = [(n, C.c_double) for n in _PARAM_DOUBLES] + [(n, C.c_int32) for n in _PARAM_INTS]

    def copy(self, **kw):
        p = SbmParams()
        C.pointer(p)[0] = self
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class StepStats(C.Structure):
    """Mirror of elsim_step_stats."""
    _fields_ = [(n, C.c_int32) for n in "it_cne it_php it_phe n_outer n_reliable status".split()] + \
               [(n, C.c_double) for n in "xrf scrnt cvgp cvgl relres_cne relres_php relres_phe seconds".split()] + \
               [("ws_used", C.c_int32), ("ws_pad_", C.c_int32), ("ws_res_plain", C.c_double),
                ("ws_res_nn", C.c_double), ("ws_alpha", C.c_double), ("ws_energy_gain", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


PROF_OP32, PROF_UPD32, PROF_DIR32, PROF_OP64, PROF_UPD64, PROF_DIR64, PROF_MG, PROF_MG_DOWN, PROF_MG_UP = range(9)
PROF_N = 9


class KernelProf(C.Structure):
    """Mirror of elsim_kernel_prof."""
    _fields_ = [("count", C.c_int32 * PROF_N), ("ms_total", C.c_double * PROF_N)]

    def mean_us(self, slot):
        return self.ms_total[slot] * 1e3 / self.count[slot] if self.count[slot] else float("nan")


_lib = None

P = C.c_void_p
I = C.c_int
D = C.c_double
IP = C.POINTER(C.c_int)
DP = C.POINTER(C.c_double)

_SIGS = {
    "elsim_init": (I, [I]),
    "elsim_last_error": (C.c_char_p, []),
    "elsim_device_count": (I, []),
    "elsim_synchronize": (I, []),
    "elsim_mem_info": (I, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "elsim_malloc": (I, [C.POINTER(P), C.c_size_t]),
    "elsim_free": (I, [P]),
    "elsim_memcpy_h2d": (I, [P, P, C.c_size_t]),
    "elsim_memcpy_d2h": (I, [P, P, C.c_size_t]),
    "elsim_memcpy_d2d": (I, [P, P, C.c_size_t]),
    "elsim_memset": (I, [P, I, C.c_size_t]),
    "elsim_timer_start": (I, []),
    "elsim_timer_stop": (I, [C.POINTER(C.c_float)]),
    "elsim_sbm_default_params": (None, [C.POINTER(SbmParams)]),
    "elsim_nogrdbc3d": (I, [IP, IP, IP, P]),
    "elsim_absgrd3duw_xyz": (I, [IP, IP, IP, P, P, P]),
    "elsim_redist3d_xyz": (I, [IP, IP, IP, IP, P, P, DP]),
    "elsim_dev_redist3d_f32": (I, [I, I, I, I, P, P, C.c_float]),
    "elsim_dev_redist3d_f64": (I, [I, I, I, I, P, P, D]),
    "elsim_ls_frontend": (I, [IP, IP, IP, P, IP, IP, IP, IP, DP, IP, P, P]),
    "elsim_refine2_sdf": (I, [IP, IP, IP, P, P]),
    "elsim_dev_refine2_sdf": (I, [I, I, I, P, P]),
    "elsim_rm_phase": (I, [IP, IP, IP, P, P, IP, P]),
    "elsim_dev_rm_phase": (I, [I, I, I, P, P, I, P]),
    "elsim_dev_ls_frontend": (I, [I, I, I, P, I, I, I, I, D, I, P, P]),
    "elsim_accdiff_phase": (I, [IP, IP, IP, P, IP, IP, P]),
    "elsim_dev_accdiff_phase": (I, [I, I, I, P, I, I, P]),
    "elsim_label_connected": (I, [IP, IP, IP, P, IP, P, P]),
    "elsim_dev_label_connected": (I, [I, I, I, P, I, P, P]),
    "elsim_rm_threshold": (I, [IP, IP, IP, P, P, P]),
    "elsim_tbl_chk_pln_3da": (I, [IP, IP, IP, IP, P, P, P, P]),
    "elsim_sbm_create": (I, [IP, IP, IP, C.POINTER(SbmParams), P, P, IP, P, C.POINTER(C.c_int64)]),
    "elsim_sbm_destroy": (I, [C.POINTER(C.c_int64)]),
    "elsim_sbm_get": (I, [C.POINTER(C.c_int64), IP, P]),
    "elsim_sbm_set": (I, [C.POINTER(C.c_int64), IP, P]),
    "elsim_sbm_field_ptr": (I, [C.POINTER(C.c_int64), IP, C.POINTER(P)]),
    "elsim_sbm_scalars": (I, [C.POINTER(C.c_int64), DP, DP, DP, DP, IP]),
    "elsim_sbm_set_params": (I, [C.POINTER(C.c_int64), C.POINTER(SbmParams)]),
    "elsim_sbm_step": (I, [C.POINTER(C.c_int64), IP, C.POINTER(StepStats)]),
    "elsim_sbm_ch_update": (I, [C.POINTER(C.c_int64)]),
    "elsim_sbm_time_control": (I, [DP, DP, DP, DP, DP, DP]),
    "elsim_sbm_matvec": (I, [C.POINTER(C.c_int64), IP, P, P]),
    "elsim_dev_matvec_f64": (I, [I, I, I, I, P, P, P, D, P, P]),
    "elsim_dev_matvec_stored_f64": (I, [I, I, I, I, P, P, P, D, P]),
    "elsim_prof_enable": (I, [I]),
    "elsim_nn_param_count": (I, []),
    "elsim_nn_create": (I, [C.POINTER(C.c_int64)]),
    "elsim_nn_destroy": (I, [C.POINTER(C.c_int64)]),
    "elsim_nn_set_params": (I, [C.POINTER(C.c_int64), P]),
    "elsim_nn_get_params": (I, [C.POINTER(C.c_int64), P]),
    "elsim_nn_get_grads": (I, [C.POINTER(C.c_int64), P]),
    "elsim_nn_forward": (I, [C.POINTER(C.c_int64), I, I, I, P, P]),
    "elsim_nn_forward_used": (I, [C.POINTER(C.c_int64), I, I, I, P, I, P]),
    "elsim_nn_loss_grad": (I, [C.POINTER(C.c_int64), I, I, I, P, P, P, DP]),
    "elsim_nn_adam": (I, [C.POINTER(C.c_int64), C.c_float]),
    "elsim_nn_train_forward": (I, [C.POINTER(C.c_int64), I, I, I, P, P]),
    "elsim_nn_energy_loss": (I, [C.POINTER(C.c_int64), P, P, P, DP]),
    "elsim_nn_backward": (I, [C.POINTER(C.c_int64), P]),
    "elsim_nn_conv3": (I, [I, I, I, P, I, P, P, I, I, P]),
    "elsim_sbm_set_warm_start": (I, [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "elsim_sbm_ws_record": (I, [C.POINTER(C.c_int64), P, P]),
    "elsim_nn_conv3_bwd_data": (I, [I, I, I, P, I, P, I, P]),
    "elsim_nn_conv3_bwd_weight": (I, [I, I, I, P, I, P, I, P]),
    "elsim_prof_read": (I, [C.POINTER(KernelProf)]),
    "elsim_dev_matvec_f32": (I, [I, I, I, I, P, P, P, C.c_float, P, P]),
    "elsim_dev_mg_apply": (I, [I, I, I, I, P, P, D, P, P]),
    "elsim_mg_config": (I, [D, D, D]),
    "elsim_mg_config_get": (I, [DP, DP, DP]),
    "elsim_gen_sphere_pack": (I, [I, I, I, C.c_uint64, D, D, D, D, I, D, P, IP, DP]),
    "elsim_gen_ellipsoid_pack": (I, [I, I, I, C.c_uint64, D, D, D, D, D, D, P, IP, DP]),
    "elsim_gen_tunnel": (I, [I, I, I, C.c_uint64, D, P]),
    "elsim_gen_tables": (I, [I, P]),
    "elsim_solve_pot": (I, [IP, IP, IP, DP, P, P, P, IP, IP, IP, DP, IP, IP, DP]),
    "elsim_dev_solve_pot": (I, [I, I, I, D, P, P, P, I, I, I, D, I, IP, DP]),
    "elsim_dev_solve_pot_noalloc": (I, [I, I, I, D, P, P, P, I, I, I, D, I, IP, DP]),
}


def lib():
    """The loaded library (raises ElsimError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ElsimError(f"{LIB_PATH} not built: run `python -m elsim_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what="elsim"):
    """Raise on a negative status; return the status otherwise."""
    if rc < 0:
        msg = lib().elsim_last_error()
        raise ElsimError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
    return rc


def iref(v):
    return C.byref(C.c_int(int(v)))


def ptr(a):
    """Host pointer of a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def exported_symbols():
    """Names declared in include/elsim.h that the library must export."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "elsim.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(elsim_[a-z0-9_]+)\s*\(", txt)))


class DeviceArray:
    """Owning device buffer with numpy-style shape/dtype (no torch)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if np.iterable(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().elsim_malloc(C.byref(p), max(self.nbytes, 1)), "elsim_malloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(lib().elsim_memcpy_h2d(d.ptr, ptr(a), a.nbytes), "h2d")
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().elsim_memcpy_d2h(ptr(out), self.ptr, self.nbytes), "d2h")
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().elsim_memcpy_h2d(self.ptr, ptr(a), a.nbytes), "h2d")

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().elsim_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
