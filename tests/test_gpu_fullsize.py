"""The 256^3 electrochemical-step problem of the benchmark itself, checked
through properties that do not depend on size: the FP64 operator bit-exact
against the oracle on the full grid, the first step of the mixed-precision
solver against the FP64 Jacobi-PCG solver from the same state (every solve to
relres 1e-8), and a bit-identical rerun of the mixed step."""
import numpy as np
import pytest

import elsim_amd as E
from elsim_amd._lib import F_CNE, F_CNP, F_PHE, F_PHP

pytestmark = pytest.mark.gpu
N = 256
FIELDS = (F_CNP, F_CNE, F_PHP, F_PHE)


@pytest.fixture(scope="module")
def sdf():
    # the benchmark's packing: seed 1, radii 6-16, solid fraction target ~U[0.40, 0.60]
    target = float(np.random.default_rng(1).uniform(0.40, 0.60))
    s, _, _ = E.sphere_pack(N, 1, 6.0, 16.0, target=target, hardcore=0.85, connect=True)
    return s


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_operator_256_bit_exact(gpu, oracle_lib, sdf):
    O = oracle_lib
    rng = np.random.default_rng(2)
    psi = 0.5 * (1.0 + np.tanh(np.pad(sdf.astype(np.float64), 1, mode="edge") / 2.0))
    kap = 3.3 * psi + 1.0e-7
    p = rng.standard_normal(kap.shape)
    ref = O.apply_op(E.OP_PHP, kap, None, O.prj_bc(E.OP_PHP, p), 1e-3, 1.0)
    dp, dk = E.DeviceArray.from_host(O.flat(p)), E.DeviceArray.from_host(O.flat(kap))
    out = E.DeviceArray(N ** 3, np.float64)
    E.check(gpu.elsim_dev_matvec_f64(E.OP_PHP, N, N, N, dp.ptr, dk.ptr, None, 1.0, out.ptr, None), "matvec")
    assert np.array_equal(O.unflat(out.to_host(), (N, N, N)), ref)
    for b in (dp, dk, out):
        b.free()


def run_step(sdf, mode):
    s = E.Sbm(sdf, -sdf, E.synthetic_params(), E.tables())
    s.set(F_CNP, E.initial_cnp(N, 1))
    st = s.step(mode)
    out = [interior(s.get(f)) for f in FIELDS]
    s.close()
    return st, out


def test_first_step_256_mixed_vs_fp64(gpu, sdf):
    st_m, mixed = run_step(sdf, E.MODE_MIXED)
    st_d, fp64 = run_step(sdf, E.MODE_FP64_PCG)
    for st in (st_m, st_d):
        assert st.status == 0, st.as_dict()
        assert st.relres_php <= 1e-8 and st.relres_phe <= 1e-8, st.as_dict()
    for f, a, b in zip(FIELDS, mixed, fp64):
        assert rel(a, b) <= 1e-6, (f, rel(a, b))
    # a rerun of the mixed step is bit-identical (fixed-order reductions)
    st_r, again = run_step(sdf, E.MODE_MIXED)
    assert st_r.it_php == st_m.it_php and st_r.it_phe == st_m.it_phe
    for a, b in zip(mixed, again):
        assert np.array_equal(a, b)
