"""Argument errors of the C-ABI are reported before any device call, so these
run without a GPU: bad grid sizes, face conditions, modes, tolerances, null
pointers and unknown handles all return ELSIM_ERR_ARG with a message."""
import ctypes as C

import numpy as np
import pytest

import elsim_amd as E
from elsim_amd._lib import ERR_ARG, iref, ptr


def solve(n=(8, 8, 8), bc=(1, 1), mode=E.MODE_FP64, rtol=1e-8, maxit=10, dx=1.0, null=False):
    L = E.lib()
    n1, n2, n3 = n
    g = max((n1 + 2) * (n2 + 2) * (n3 + 2), 1)
    kap, phi, rhs = np.ones(g), np.zeros(g), np.zeros(max(n1 * n2 * n3, 1))
    it, rr = C.c_int(), C.c_double()
    return L.elsim_solve_pot(iref(n1), iref(n2), iref(n3), C.byref(C.c_double(dx)), None if null else ptr(kap),
                             ptr(rhs), ptr(phi), iref(bc[0]), iref(bc[1]), iref(mode), C.byref(C.c_double(rtol)),
                             iref(maxit), C.byref(it), C.byref(rr))


@pytest.mark.parametrize("kw, msg", [
    (dict(n=(0, 8, 8)), "bad grid"),
    (dict(n=(8, -1, 8)), "bad grid"),
    (dict(bc=(0, 0)), "at least one Dirichlet"),
    (dict(bc=(2, 1)), "d2 face conditions"),
    (dict(mode=7), "unknown mode"),
    (dict(rtol=0.0), "need rtol > 0"),
    (dict(maxit=0), "maxit >= 1"),
    (dict(dx=-1.0), "dx > 0"),
    (dict(null=True), "null array"),
])
def test_solve_pot_rejects_bad_arguments(kw, msg):
    assert solve(**kw) == ERR_ARG
    assert msg in E.lib().elsim_last_error().decode()


def test_unknown_handles_are_rejected():
    L = E.lib()
    bogus = C.c_int64(987654321)
    st = E.StepStats()
    assert L.elsim_sbm_step(C.byref(bogus), iref(E.MODE_MIXED), C.byref(st)) == ERR_ARG
    assert L.elsim_sbm_destroy(C.byref(bogus)) == ERR_ARG
    assert L.elsim_sbm_set_warm_start(C.byref(bogus), None) == ERR_ARG
    assert L.elsim_nn_destroy(C.byref(bogus)) == ERR_ARG
    assert L.elsim_nn_set_params(C.byref(bogus), ptr(np.zeros(4, np.float32))) == ERR_ARG
    # a handle that looks like a pointer is not dereferenced
    fake = C.c_int64(id(st))
    assert L.elsim_nn_get_params(C.byref(fake), ptr(np.zeros(4, np.float32))) == ERR_ARG
