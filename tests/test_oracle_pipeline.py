"""Pins of the oracle's RM_UCNCT_3D sign block (RM_UCNCT_3D.f90:64-96,
115-136) and of the 2x refinement between RM_UCNCT_3D and the SBM program.
The refinement has no counterpart in the reference (its input files for SBM
were made by a step that is not in the repository): these checks fix the
stated design -- trilinear interpolation at the fine cell centres, Neumann
clamping, distances doubled -- independently of the restatement."""
import numpy as np
import pytest

import oracle as O

R4_1EM4 = float(np.float32(1.0e-4))


def interp_matrix(n):
    """(2n x n) 1-D interpolation of the refinement: fine I at coarse
    coordinate (I + 0.5)/2 (1-based), linear between the two nearest coarse
    centres, clamped to 1..n."""
    W = np.zeros((2 * n, n))
    for I in range(1, 2 * n + 1):
        x = (I + 0.5) / 2.0
        c0 = int(np.floor(x))
        w1 = x - c0
        for c, w in ((c0, 1.0 - w1), (c0 + 1, w1)):
            W[I - 1, min(max(c, 1), n) - 1] += w
    return W


@pytest.mark.parametrize("shape", [(5, 7, 3), (1, 4, 6), (9, 2, 8)])
def test_refine2_matches_separable_model(oracle_lib, shape):
    rng = np.random.default_rng(sum(shape))
    sdf = rng.uniform(-20, 20, shape)
    got = O.refine2(sdf).astype(np.float64)
    W1, W2, W3 = (interp_matrix(n) for n in shape)
    want = 2.0 * np.einsum("ai,bj,ck,ijk->abc", W1, W2, W3, sdf)
    assert got.shape == tuple(2 * n for n in shape)
    assert np.allclose(got, want, rtol=2e-7, atol=1e-6)


def test_refine2_exact_on_affine_fields(oracle_lib):
    """Away from the clamped faces trilinear interpolation reproduces affine
    fields; with dyadic coefficients every product and sum is exact."""
    n1, n2, n3 = 6, 5, 7
    i, j, k = np.meshgrid(np.arange(1, n1 + 1), np.arange(1, n2 + 1), np.arange(1, n3 + 1), indexing="ij")
    a, b, c, e = 0.5, -0.25, 2.0, 1.125
    sdf = a * i + b * j + c * k + e
    got = O.refine2(sdf).astype(np.float64)
    I, J, K = np.meshgrid(np.arange(1, 2 * n1 + 1), np.arange(1, 2 * n2 + 1), np.arange(1, 2 * n3 + 1), indexing="ij")
    want = 2.0 * (a * (I + 0.5) / 2 + b * (J + 0.5) / 2 + c * (K + 0.5) / 2 + e)
    inner = (slice(1, -1),) * 3  # the first and last fine cells sit outside the coarse centres
    assert np.array_equal(got[inner], want[inner])


def test_refine2_constant_and_faces(oracle_lib):
    sdf = np.full((4, 3, 5), -3.5)
    assert np.all(O.refine2(sdf) == np.float32(-7.0))
    # a field varying along d1 only, dyadic values (every operation exact):
    # fine cell 1 lies outside the first coarse centre and takes the clamped
    # value, fine cell 2 blends 3/4 of coarse 1 with 1/4 of coarse 2, ...
    v = np.array([1.5, -2.25, 0.75, 4.0])
    sdf = np.broadcast_to(v[:, None, None], (4, 3, 5)).copy()
    got = O.refine2(sdf)[:, 2, 3].astype(np.float64)
    want = 2.0 * np.array([v[0], 0.75 * v[0] + 0.25 * v[1], 0.25 * v[0] + 0.75 * v[1], 0.75 * v[1] + 0.25 * v[2],
                           0.25 * v[1] + 0.75 * v[2], 0.75 * v[2] + 0.25 * v[3], 0.25 * v[2] + 0.75 * v[3], v[3]])
    assert np.array_equal(got, want)


def test_rm_phase_matches_numpy(oracle_lib):
    rng = np.random.default_rng(4)
    shape = (9, 11, 6)
    rd = rng.uniform(-4, 4, shape)
    rd[np.abs(rd) < 1e-3] = 0.5  # keep 0.5(1 + tanh) away from its 0.5 threshold
    cn = np.where(rng.random(shape) < 0.3, 0.0, rng.uniform(0, 1, shape))
    cn[0, 0, 0] = R4_1EM4  # at the threshold: kept (the reference zeroes Cn < 1e-4)
    for phase in (0, 1):
        ps = 0.5 * (1.0 + np.tanh(rd / 1.0))
        if phase == 1:
            ps = 1.0 - ps
        ps = np.where(cn < R4_1EM4, 0.0, ps)
        want = np.where(np.pad(ps, 1, mode="edge") > 0.5, 1.0, -1.0)
        got = O.rm_phase(rd, cn, phase)
        assert np.array_equal(got, want)
    # the two phases are complementary where the voxel is connected
    s0, s1 = O.rm_phase(rd, np.ones(shape), 0), O.rm_phase(rd, np.ones(shape), 1)
    assert np.array_equal(s0, -s1)
