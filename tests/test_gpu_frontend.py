"""GPU parity of LevelSet_3D's geometry front-end (LevelSet_3D.f90:84-144)
and of a whole frame (front-end + ReDist3D_XYZ(175), LS:146-148) against the
oracle: bit-exact (the block has FP64 arithmetic in a fixed order and IEEE
square roots only)."""
import numpy as np
import pytest

import oracle as O
from elsim_amd import frontend as F

pytestmark = pytest.mark.gpu


def voxels(shape, seed):
    return (np.random.default_rng(seed).random(shape) > 0.45).astype(np.float64)


# (image yo, xo, zo), (py, pz, sep, blocks), tr, n_ac: odd sizes, separators
# of 0..7 voxels, 1..4 thickness copies, tunnels whose loop offsets wrap
# into later planes of TmpB, no tunnel, and a tunnel wider than the grid
CASES = [
    ((26, 31, 9), (17, 23, 5, 4), 3.5, 0),
    ((26, 31, 9), (17, 23, 5, 4), 7.25, 7),
    ((40, 64, 11), (33, 50, 7, 4), 9.5, 50),
    ((21, 19, 13), (21, 17, 0, 1), 0.0, 12),
    ((30, 45, 6), (29, 35, 3, 3), 40.0, 3),
]


@pytest.mark.parametrize("case", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_frontend_bit_exact(gpu, oracle_lib, case):
    shape, (py, pz, sep, nb), tr, n_ac = case
    zo = shape[2]
    grid = (py, sep + nb * zo, pz)
    tmpa = voxels(shape, 7 + n_ac)
    sg_o, ds_o = O.ls_frontend(tmpa, grid, sep, tr, n_ac)
    sg_g, ds_g = F.ls_frontend(tmpa, grid, sep, tr, n_ac)
    assert np.array_equal(ds_g, ds_o)
    assert np.array_equal(sg_g, sg_o)


def test_frame_with_reinit_bit_exact(gpu, oracle_lib):
    """Front-end, then dsF = 0 and 175 reinitialisation sweeps with dt = 0.1d0,
    as one frame of the reference loop writes it."""
    shape, grid, sep, tr = (36, 40, 10), (31, 6 + 4 * 10, 33), 6, F.TURD[3] / 3
    tmpa = voxels(shape, 3)
    sg_o, _ = O.ls_frontend(tmpa, grid, sep, tr, 50)
    f_o = O.redist3d(175, sg_o, np.zeros_like(sg_o), 0.1)
    sg_g, f_g = F.ls_frame(tmpa, grid, sep, tr)
    assert np.array_equal(sg_g, sg_o)
    assert np.array_equal(f_g, f_o)


def test_rejects_geometry_outside_the_image(gpu):
    import elsim_amd as E
    with pytest.raises(E.ElsimError):
        F.ls_frontend(np.ones((6, 4, 3)), (5, 2 + 4 * 3, 4), 2, 1.0, 0)
