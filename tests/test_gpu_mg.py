"""The multigrid-preconditioned mixed solver (ELSIM_MODE_MIXED_MG) on the GPU.

* One V-cycle (elsim_dev_mg_apply) against the FP64 numpy model of the same
  cycle (tests/mg_model.py) on odd and even boxes, for every operator: the
  Galerkin hierarchy, the smoothing, the coarse solve and the signs.  The GPU
  cycle runs in FP32 with an FP32 coefficient, so it agrees with the model to
  FP32 rounding.
* The GPU cycle is symmetric (u.M v = v.M u to FP32 rounding), as CG needs.
* A potential solve with it needs far fewer iterations than Jacobi-PCG and
  still stops on the FP64 true residual at 1e-8.
The solves' parity with the oracle (steps, the 64^3 solve, face rules) is in
test_gpu_sbm.py, test_gpu_solve_pot.py and test_gpu_sbm_reference.py, which
run this mode beside the others.
"""
import numpy as np
import pytest

import elsim_amd as E
import mg_model as M

pytestmark = pytest.mark.gpu


def dev_vcycle(op, kap, pse, scale, r):
    g = kap.shape
    dk = E.DeviceArray.from_host(np.asfortranarray(kap).ravel(order="F"))
    dp = E.DeviceArray.from_host(np.asfortranarray(pse).ravel(order="F")) if pse is not None else None
    rg = np.zeros(g, np.float32)
    rg[1:-1, 1:-1, 1:-1] = r
    dr = E.DeviceArray.from_host(np.asfortranarray(rg).ravel(order="F"))
    dz = E.DeviceArray(int(np.prod(g)), np.float32)
    n1, n2, n3 = (s - 2 for s in g)
    E.check(E.lib().elsim_dev_mg_apply(op, n1, n2, n3, dk.ptr, dp.ptr if dp else None, scale, dr.ptr, dz.ptr),
            "mg_apply")
    z = dz.to_host().reshape(g, order="F")
    assert not np.any(z[0]) and not np.any(z[-1]) and not np.any(z[:, 0]) and not np.any(z[:, :, -1])
    return z[1:-1, 1:-1, 1:-1].astype(np.float64)


def coefficient(shape, seed, contrast):
    """Ghosted coefficient: uniform in [0.1, 1] (contrast False) or a
    psi-weighted sphere pack, psi = 0.5(1+tanh(d/2)) + 1e-7 (contrast True)."""
    rng = np.random.default_rng(seed)
    if not contrast:
        return np.asfortranarray(rng.uniform(0.1, 1.0, tuple(s + 2 for s in shape)))
    sdf, _, _ = E.sphere_pack(shape, seed, 3.0, 6.0, target=0.5, hardcore=0.85, connect=True)
    psi = 0.5 * (1.0 + np.tanh(sdf.astype(np.float64) / 2.0)) + 1e-7
    return np.asfortranarray(np.pad(psi, 1, mode="edge"))


CASES = [((20, 12, 10), False), ((33, 17, 9), False), ((24, 24, 24), True), ((64, 32, 48), True),
         ((5, 3, 2), False)]


@pytest.mark.parametrize("op", [M.OP_PHP, M.OP_PHE, M.OP_POT_DD, M.OP_CNE])
@pytest.mark.parametrize("shape,contrast", CASES)
def test_vcycle_matches_model(gpu, op, shape, contrast):
    kap = coefficient(shape, 3, contrast)
    rng = np.random.default_rng(11)
    pse = np.asfortranarray(rng.uniform(0.2, 1.0, kap.shape)) if op == M.OP_CNE else None
    scale = 1e-3 if op == M.OP_CNE else 1.0
    r = rng.standard_normal(shape).astype(np.float32)
    z = dev_vcycle(op, kap, pse, scale, r)
    ref = M.VCycle(op, kap, pse, scale)(r.astype(np.float64))
    e = np.linalg.norm(z - ref) / np.linalg.norm(ref)
    print(f"op {op} shape {shape} contrast {contrast}: levels {[l[1].shape for l in M.VCycle(op, kap, pse, scale).levels]}"
          f" rel err {e:.2e}")
    # FP32 cycle against the FP64 model: a few FP32 roundings (measured
    # 4e-8 .. 2.3e-7 over these cases, high contrast included); an indexing
    # or sign mistake anywhere in the hierarchy gives O(1)
    assert e <= 2e-6, e


@pytest.mark.parametrize("op", [M.OP_PHP, M.OP_CNE])
def test_vcycle_symmetric(gpu, op):
    shape = (40, 24, 16)
    kap = coefficient(shape, 5, True)
    rng = np.random.default_rng(2)
    pse = np.asfortranarray(rng.uniform(0.2, 1.0, kap.shape)) if op == M.OP_CNE else None
    scale = 1e-3 if op == M.OP_CNE else 1.0
    u = rng.standard_normal(shape).astype(np.float32)
    v = rng.standard_normal(shape).astype(np.float32)
    mu, mv = dev_vcycle(op, kap, pse, scale, u), dev_vcycle(op, kap, pse, scale, v)
    a, b = float(np.vdot(v, mu)), float(np.vdot(u, mv))
    assert abs(a - b) <= 1e-5 * np.linalg.norm(mu) * np.linalg.norm(v), (a, b)
    # definite with the operator's sign: negative for the potentials
    sgn = 1.0 if op == M.OP_CNE else -1.0
    assert sgn * float(np.vdot(u, mu)) > 0


def test_solve_needs_fewer_iterations(gpu):
    kap, rhs, phi, _ = E.potential_problem(64, 2)
    x_j, rc_j, it_j, rr_j = E.solve_pot(kap, rhs, phi, mode=E.MODE_MIXED)
    x_m, rc_m, it_m, rr_m = E.solve_pot(kap, rhs, phi, mode=E.MODE_MIXED_MG)
    print(f"64^3 potential solve: Jacobi-PCG {it_j} iterations, multigrid-PCG {it_m}")
    assert rc_j == 0 and rc_m == 0 and rr_j <= 1e-8 and rr_m <= 1e-8
    assert it_m * 5 <= it_j
    e = np.linalg.norm(x_m - x_j) / np.linalg.norm(x_j)
    assert e <= 1e-6, e
