"""One potential solve on its own (elsim_solve_pot / SBM_Solve_Pot; the
reference's inline solve block SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3.f90:657-714)
on the GPU against the FP64 oracle:

* the 64^3 plain-FP64 configuration (non-overlapping spheres r 4-8,
  psi = 0.5(1+tanh(d/2)), kappa = psi + 1e-7, Dirichlet 0 / 1 on the d2 faces,
  rhs = |grad psi| j_BV) with the reference CG, the FP64 Jacobi-PCG and the
  mixed-precision solver, fields to 1e-6 relative L2 at relres 1e-8;
* the same from Fortran through SBM_Solve_Pot;
* the closed form of a uniform conductor (a linear profile between the
  ghost-plane values), and each single-Dirichlet face rule;
* failure paths: an iteration cap returns ELSIM_NOT_CONVERGED (and ends the
  Fortran program through ERROR STOP), a non-finite right-hand side returns
  ELSIM_ERR_NAN instead of a converged status."""
import os
import subprocess

import numpy as np
import pytest

import elsim_amd as E
from elsim_amd import fileio

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def c1():
    return E.potential_problem(64, 1)


@pytest.fixture(scope="module")
def c1_oracle(c1, oracle_lib):
    kap, rhs, phi, _ = c1
    x, st, it, rr = oracle_lib.solve_pot(kap, rhs, phi)
    assert st == 0 and rr <= 1e-8
    return x, it


@pytest.mark.parametrize("mode", [E.MODE_FP64, E.MODE_FP64_PCG, E.MODE_MIXED, E.MODE_MIXED_MG])
def test_c1_64_matches_oracle(gpu, c1, c1_oracle, mode):
    kap, rhs, phi, info = c1
    ref, it_ref = c1_oracle
    x, rc, it, rr = E.solve_pot(kap, rhs, phi, mode=mode)
    print(f"mode {mode}: {it} iterations (oracle {it_ref}), relres {rr:.2e}, solid {info['solid_fraction']:.3f}")
    assert rc == 0 and rr <= 1e-8
    assert rel(x[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1]) <= 1e-6
    # the Dirichlet ghost planes keep their values, the others are mirrors
    assert np.all(x[:, 0, :] == 0.0) and np.all(x[:, -1, :] == 1.0)
    assert np.array_equal(x[0, 1:-1, :], x[1, 1:-1, :]) and np.array_equal(x[:, 1:-1, -1], x[:, 1:-1, -2])
    if mode == E.MODE_FP64:
        # the same algorithm; reductions differ only in summation order
        assert abs(it - it_ref) <= max(2, it_ref // 50)


def test_c1_64_from_fortran(gpu, c1, c1_oracle, tmp_path):
    kap, rhs, phi, _ = c1
    ref, _ = c1_oracle
    for name, a in (("kap", kap), ("rhs", rhs), ("phi", phi)):
        fileio.write_record(tmp_path / f"{name}.dat", a)
    exe = os.path.join(os.path.dirname(E.__file__), "lib", "solve_check")
    r = subprocess.run([exe, str(tmp_path), "64", "64", "64", "1", "1", str(E.MODE_FP64), "100000"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    x = fileio.read_record(tmp_path / "phi_out.dat", phi.shape, np.float64)
    it, rr = open(tmp_path / "solve.txt").read().split()
    assert float(rr) <= 1e-8
    assert rel(x[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1]) <= 1e-6


def test_uniform_conductor_is_linear(gpu):
    n1, n2, n3 = 9, 21, 7
    kap = np.full((n1 + 2, n2 + 2, n3 + 2), 2.5)
    phi = np.zeros_like(kap)
    phi[:, -1, :] = 1.0
    x, rc, _, rr = E.solve_pot(kap, np.zeros((n1, n2, n3)), phi, mode=E.MODE_FP64)
    assert rc == 0
    # ghost-centred Dirichlet values 0 at j = 0 and 1 at j = n2 + 1
    want = np.arange(1, n2 + 1) / (n2 + 1.0)
    assert np.allclose(x[1:-1, 1:-1, 1:-1], want[None, :, None], rtol=0, atol=1e-9)


@pytest.mark.parametrize("bc", [(0, 1), (1, 0)])
def test_single_dirichlet_face_matches_oracle(gpu, oracle_lib, bc):
    kap, rhs, phi, _ = E.potential_problem(24, 3, rmin=3.0, rmax=6.0)
    ref, st, _, _ = oracle_lib.solve_pot(kap, rhs, phi, *bc)
    assert st == 0
    for mode in (E.MODE_FP64, E.MODE_MIXED, E.MODE_MIXED_MG):
        x, rc, _, rr = E.solve_pot(kap, rhs, phi, *bc, mode=mode)
        assert rc == 0 and rr <= 1e-8
        assert rel(x[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1]) <= 1e-6


def test_iteration_cap_reports_not_converged(gpu, c1, tmp_path):
    kap, rhs, phi, _ = c1
    for mode in (E.MODE_FP64, E.MODE_FP64_PCG, E.MODE_MIXED, E.MODE_MIXED_MG):
        _, rc, it, rr = E.solve_pot(kap, rhs, phi, mode=mode, maxit=1)
        assert rc == E._lib.NOT_CONVERGED, (mode, rc)
        assert it <= 2 and rr > 1e-8
    # the Fortran wrapper ends the program (the reference would continue silently, SBM:710)
    for name, a in (("kap", kap), ("rhs", rhs), ("phi", phi)):
        fileio.write_record(tmp_path / f"{name}.dat", a)
    exe = os.path.join(os.path.dirname(E.__file__), "lib", "solve_check")
    r = subprocess.run([exe, str(tmp_path), "64", "64", "64", "1", "1", str(E.MODE_FP64), "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "did not converge" in r.stderr
    assert not (tmp_path / "phi_out.dat").exists()


def test_nonfinite_rhs_is_an_error(gpu):
    kap, rhs, phi, _ = E.potential_problem(24, 3, rmin=3.0, rmax=6.0)
    rhs = rhs.copy()
    rhs[3, 4, 5] = np.nan
    for mode in (E.MODE_FP64, E.MODE_FP64_PCG, E.MODE_MIXED, E.MODE_MIXED_MG):
        with pytest.raises(E.ElsimError, match=r"\(-3\)"):
            E.solve_pot(kap, rhs, phi, mode=mode)
    # the library stays usable after the error
    _, rc, _, rr = E.solve_pot(kap, np.zeros_like(rhs), phi, mode=E.MODE_MIXED)
    assert rc == 0 and rr <= 1e-8
