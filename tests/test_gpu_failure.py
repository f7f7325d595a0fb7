"""Failure handling of the learned warm start on the GPU: a network whose
output is not finite changes nothing (the Galerkin step along it is zero, so
the run is bit-identical to the extrapolated start), and destroying an
attached network detaches it instead of leaving the state a dangling
pointer."""
import ctypes as C

import numpy as np
import pytest

import elsim_amd as E
from elsim_amd import nn as NN
from elsim_amd._lib import ERR_ARG, F_CNE, F_CNP, F_PHE, F_PHP

pytestmark = pytest.mark.gpu
N = 32
STEPS = 6


def state(ws_mode):
    sdf, _, _ = E.sphere_pack(N, 21, 4.0, 8.0, target=0.5, hardcore=0.85, connect=True)
    s = E.Sbm(sdf, -sdf, E.synthetic_params(ws_mode=ws_mode), E.tables())
    s.set(F_CNP, E.initial_cnp(N, 21))
    return s


def run(s):
    st = [s.step(E.MODE_MIXED) for _ in range(STEPS)]
    return st, [s.get(f) for f in (F_CNP, F_CNE, F_PHP, F_PHE)]


def test_nonfinite_network_output_is_ignored(gpu):
    ref = state(E.WS_EXTRAP)
    st_ref, f_ref = run(ref)
    ref.close()
    net = NN.WarmStartNet()
    net.set_params(np.full(NN.param_count(), np.nan, np.float32))
    s = state(E.WS_LEARNED)
    s.set_warm_start(net)
    st, f = run(s)
    for a in st:
        assert a.status == 0 and a.ws_alpha == 0.0 and a.ws_used <= 1
    for a, b in zip(f, f_ref):
        assert np.array_equal(a, b)
    assert [a.it_php for a in st] == [b.it_php for b in st_ref]
    s.close()
    net.close()


def test_destroying_an_attached_network_detaches_it(gpu):
    net = NN.WarmStartNet.load(NN.DEFAULT_WEIGHTS)
    s = state(E.WS_LEARNED)
    s.set_warm_start(net)
    st0 = [s.step(E.MODE_MIXED) for _ in range(3)]
    assert st0[-1].ws_used == 2
    h = C.c_int64(net.h.value)
    net.close()  # elsim_nn_destroy while attached
    for _ in range(2):
        st = s.step(E.MODE_MIXED)
        assert st.status == 0 and st.ws_used <= 1
    # the stale handle is refused
    assert E.lib().elsim_sbm_set_warm_start(C.byref(s.h), C.byref(h)) == ERR_ARG
    s.close()
