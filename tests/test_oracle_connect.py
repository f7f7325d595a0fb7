"""Pins of the oracle's connectivity path (AccDiff_3D.f90:83-243,
RM_UCNCT_3D.f90:79-82): the labeller against scipy's independent labelling,
the converged AD mask against reachability, and AD's monotone bounds."""
import numpy as np
import pytest
import scipy.ndimage as ndi

import elsim_amd as E
import oracle as O

THR = float(np.float32(0.4))


def case(shape, seed):
    sdf, _, _ = E.sphere_pack(shape, seed, 3.0, 6.0, target=0.55, hardcore=0.9, connect=False, dcut=12.0)
    return sdf.astype(np.float64)


def phase_psi(rd, phase):
    p = 0.5 * (1.0 + np.tanh(rd / 0.75))
    return 1.0 - p if phase == 1 else p


@pytest.mark.parametrize("phase", [0, 1])
def test_labels_vs_scipy(oracle_lib, phase):
    shape = (18, 22, 15)
    rd = case(shape, 3)
    lab, con = O.label_connected(rd, phase)
    m = phase_psi(rd, phase) >= THR
    sl, n = ndi.label(m)  # 6-connectivity by default in 3-D
    assert np.array_equal(lab >= 0, m)
    # same partition; labels are the smallest Fortran-order index per component
    flat_idx = np.arange(m.size).reshape(shape, order="F")
    for c in range(1, n + 1):
        sel = sl == c
        assert np.all(lab[sel] == flat_idx[sel].min())
    face = sl[:, -1, :] if phase == 0 else sl[:, 0, :]
    touching = set(np.unique(face)) - {0}
    assert np.array_equal(con.astype(bool), np.isin(sl, list(touching)))


@pytest.mark.parametrize("phase", [0, 1])
def test_converged_accdiff_equals_reachability(oracle_lib, phase):
    shape = (16, 20, 12)
    rd = case(shape, 9)
    cn = O.accdiff_phase(rd, phase, 1501)
    _, con = O.label_connected(rd, phase)
    assert np.array_equal(cn >= np.float32(1e-4), con.astype(bool))


def test_accdiff_bounded_and_partial_front(oracle_lib):
    shape = (10, 60, 10)
    rd = np.full(shape, 3.0)  # all solid: an open slab
    cn = O.accdiff_phase(rd, 0, 51)
    assert cn.min() >= 0.0 and cn.max() <= 1.0
    # the snapped front advances about 0.6 voxel per iteration from the
    # Dirichlet face (high d2 face for the solid)
    snapped = np.where((cn == 1.0).all(axis=(0, 2)))[0]
    reach = shape[1] - snapped.min()
    assert 20 <= reach <= 40
