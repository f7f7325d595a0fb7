"""The allocation-free potential solve (elsim_dev_solve_pot_noalloc) against
the allocating one (elsim_dev_solve_pot) through the C-ABI: identical fields,
iteration counts and residuals, on two calls in a row (the second on a
smaller grid, so the shared work space holds the first call's data), for the
FP64 reference CG and the mixed-precision solver; grids above the stated
maximum are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import elsim_amd as E
from elsim_amd._lib import ERR_ARG, DeviceArray

pytestmark = pytest.mark.gpu
MAX_G = 287496


def solve(entry, kap, rhs, phi, mode):
    n1, n2, n3 = rhs.shape
    dk = DeviceArray.from_host(np.asfortranarray(kap).ravel(order="F"))
    dr = DeviceArray.from_host(np.asfortranarray(rhs).ravel(order="F"))
    dp = DeviceArray.from_host(np.asfortranarray(phi).ravel(order="F"))
    it, rr = C.c_int(), C.c_double()
    rc = getattr(E.lib(), entry)(n1, n2, n3, C.c_double(1.0), dk.ptr, dr.ptr, dp.ptr, 1, 1, mode, C.c_double(1e-8),
                                 100000, C.byref(it), C.byref(rr))
    out = dp.to_host()
    for b in (dk, dr, dp):
        b.free()
    return rc, out, it.value, rr.value


@pytest.mark.parametrize("mode", [E.MODE_FP64, E.MODE_MIXED])
def test_noalloc_matches_allocating_solve(gpu, mode):
    big = E.potential_problem(64, 1)      # 66^3 = 287,496 ghosted elements: the maximum
    small = E.potential_problem(24, 5, rmin=3.0, rmax=6.0)
    for kap, rhs, phi, _ in (big, small):
        ref = solve("elsim_dev_solve_pot", kap, rhs, phi, mode)
        got = solve("elsim_dev_solve_pot_noalloc", kap, rhs, phi, mode)
        assert ref[0] == 0 and got[0] == 0, E.lib().elsim_last_error()
        assert ref[3] <= 1e-8
        assert got[2] == ref[2] and got[3] == ref[3]
        assert np.array_equal(got[1], ref[1])


def test_noalloc_rejects_grids_above_its_maximum(gpu):
    kap, rhs, phi, _ = E.potential_problem(64, 1)
    n = 65  # 67^3 ghosted elements > the maximum
    k2 = np.pad(kap, ((0, 1), (0, 1), (0, 1)), mode="edge")
    r2 = np.pad(rhs, ((0, 1), (0, 1), (0, 1)), mode="edge")
    p2 = np.pad(phi, ((0, 1), (0, 1), (0, 1)), mode="edge")
    assert (n + 2) ** 3 > MAX_G
    rc, _, _, _ = solve("elsim_dev_solve_pot_noalloc", k2, r2, p2, E.MODE_MIXED)
    assert rc == ERR_ARG
    assert "exceed the maximum" in E.lib().elsim_last_error().decode()


def test_noalloc_refuses_the_multigrid_mode(gpu):
    kap, rhs, phi, _ = E.potential_problem(24, 5, rmin=3.0, rmax=6.0)
    rc, _, _, _ = solve("elsim_dev_solve_pot_noalloc", kap, rhs, phi, E.MODE_MIXED_MG)
    assert rc == ERR_ARG
    assert "multigrid" in E.lib().elsim_last_error().decode()
