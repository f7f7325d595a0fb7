"""Pins of the oracle's restatement of LevelSet_3D's geometry front-end
(LevelSet_3D.f90:84-144): tunnel carving with the loop's out-of-extent
TmpB(i,j,k) indexing, thickness stacking, Allen-Cahn smoothing and the sign
function.  The reference program reads a voxel file that is not in the
repository and has no callable subroutine for this block, so these checks
are independent models of each step rather than the reference build."""
import numpy as np
import pytest

import oracle as O

R4_064 = float(np.float32(0.64))
R4_005 = float(np.float32(0.05))


def reorient(tmpa):
    """TmpB(1:yo,k,1:xo) = TmpA(1:yo,1:xo,zo-k+1) (LS:84-87)."""
    return np.transpose(tmpa[:, :, ::-1], (0, 2, 1)).copy()


def carve_bruteforce(tmpb, py, px, pz, tr):
    """LS:89-108 with every loop cell written at its column-major offset in
    TmpB, as Fortran does for TmpB(i,j,k) with j beyond TmpB's extent."""
    yo, zo, xo = tmpb.shape
    flatb = tmpb.ravel(order="F").copy()
    i, j, k = np.meshgrid(np.arange(1, py + 1), np.arange(1, px + 1), np.arange(1, pz + 1), indexing="ij")
    r1 = np.sqrt((i - 1.0) ** 2 + (k - float(pz)) ** 2)
    r2 = np.sqrt((i - float(py)) ** 2 + (k - 1.0) ** 2)
    hit = (r1 < tr) | (r2 < tr)
    off = (i - 1) + yo * (j - 1) + yo * zo * (k - 1)
    flatb[off[hit]] = 0.0
    return flatb.reshape(tmpb.shape, order="F")


def stack(tmpb, py, px, pz, sep):
    """LS:76, 110-121: zero separator, then copies as is / mirrored in d2."""
    zo = tmpb.shape[1]
    d = np.zeros((py, px, pz))
    for j in range(sep + 1, px + 1):
        r, blk = (j - sep - 1) % zo, (j - sep - 1) // zo
        jb = r + 1 if blk % 2 == 0 else zo - r
        d[:, j - 1, :] = tmpb[:py, jb - 1, :pz]
    return d


def voxels(shape, seed):
    return (np.random.default_rng(seed).random(shape) > 0.45).astype(np.float64)


@pytest.mark.parametrize("tr", [0.0, 3.5, 7.25])
def test_carve_and_stack_match_bruteforce(oracle_lib, tr):
    yo, xo, zo = 26, 31, 9
    py, pz, sep = 17, 23, 5
    px = sep + 4 * zo
    tmpa = voxels((yo, xo, zo), 1)
    _, ds = O.ls_frontend(tmpa, (py, px, pz), sep, tr, n_ac=0)
    ref = stack(carve_bruteforce(reorient(tmpa), py, px, pz, tr), py, px, pz, sep)
    assert np.array_equal(ds[1:-1, 1:-1, 1:-1], ref)


def test_tunnel_aliasing_reaches_later_planes(oracle_lib):
    """The tunnels are zeroed beyond their own planes: the loop's j runs to
    px > zo, so the offsets wrap into the next (px-1)//zo planes."""
    yo, xo, zo = 20, 40, 6
    py, pz, sep = 12, 30, 4
    px = sep + 4 * zo
    tmpa = np.ones((yo, xo, zo))
    tr = 2.5
    _, ds = O.ls_frontend(tmpa, (py, px, pz), sep, tr, n_ac=0)
    d = ds[1:-1, 1:-1, 1:-1]
    # the geometric disc of the first tunnel around (i, k) = (1, pz) reaches
    # down to k = pz - 2; the wrapped writes extend it to k = pz - 2 + 4 only
    # where it exists, so it is zero in planes pz-2 .. pz but not at k = pz - 7
    assert np.all(d[0, sep:, pz - 1] == 0.0)
    assert np.all(d[0, sep:, pz - 8] == 1.0)
    # the second tunnel around (py, 1): wrapped offsets reach planes 2..5 of
    # the rows whose loop plane is in the disc
    assert d[py - 1, sep, 0] == 0.0
    model = stack(carve_bruteforce(reorient(tmpa), py, px, pz, tr), py, px, pz, sep)
    no_wrap = np.ones_like(model)
    i, k = np.meshgrid(np.arange(1, py + 1), np.arange(1, pz + 1), indexing="ij")
    disc = (np.sqrt((i - 1.0) ** 2 + (k - float(pz)) ** 2) < tr) | (np.sqrt((i - float(py)) ** 2 + (k - 1.0) ** 2) < tr)
    no_wrap[:, sep:, :] = np.where(disc[:, None, :], 0.0, 1.0)
    no_wrap[:, :sep, :] = 0.0
    assert np.array_equal(d, model)
    assert not np.array_equal(d, no_wrap)  # the wrap changes the result


def test_stacking_mirrors_alternate_blocks(oracle_lib):
    yo, xo, zo = 14, 16, 7
    py, pz, sep = 10, 12, 3
    px = sep + 4 * zo
    tmpa = voxels((yo, xo, zo), 2)
    _, ds = O.ls_frontend(tmpa, (py, px, pz), sep, 0.0, n_ac=0)
    d = ds[1:-1, 1:-1, 1:-1]
    assert np.all(d[:, :sep, :] == 0.0)
    b0, b1 = d[:, sep:sep + zo, :], d[:, sep + zo:sep + 2 * zo, :]
    assert np.array_equal(b1, b0[:, ::-1, :])
    assert np.array_equal(d[:, sep + 2 * zo:sep + 3 * zo, :], b0)
    assert np.array_equal(d[:, sep + 3 * zo:, :], b1)


def test_allen_cahn_step_on_a_quadratic_profile(oracle_lib):
    """One iteration on d = c i^2 (constant along d2, d3 inside a block):
    the discrete Laplacian is exactly 2c away from the faces, so the new value
    is d - 0.05 (2d(1-d)(1-2d) - 0.64 * 2c) with the REAL(4) literals."""
    yo, xo, zo = 12, 20, 8
    py, pz, sep = 12, 16, 2
    px = sep + 2 * zo
    c = 1.0 / 256.0
    ii = np.arange(1, yo + 1, dtype=np.float64)
    tmpa = np.broadcast_to((c * ii * ii)[:, None, None], (yo, xo, zo)).copy()
    _, ds = O.ls_frontend(tmpa, (py, px, pz), sep, 0.0, n_ac=1)
    d = ds[1:-1, 1:-1, 1:-1]
    x = c * ii[1:py - 1] ** 2
    want = x - R4_005 * (2 * x * (1 - x) * (1 - 2 * x) - R4_064 * 2 * c)
    got = d[1:py - 1, sep + 2, 3]
    assert np.allclose(got, want, rtol=0, atol=1e-15)
    # wrong literal kind or sign would move it by far more than that
    wrong = x - 0.05 * (2 * x * (1 - x) * (1 - 2 * x) - 0.64 * 2 * c)
    assert np.abs(got - wrong).max() > 1e-11


@pytest.mark.parametrize("value", [0.0, 1.0])
def test_allen_cahn_fixed_points(oracle_lib, value):
    yo, xo, zo = 9, 12, 5
    py, pz, sep = 8, 10, 0
    px = sep + 2 * zo
    tmpa = np.full((yo, xo, zo), value)
    sg, ds = O.ls_frontend(tmpa, (py, px, pz), sep, 0.0, n_ac=20)
    assert np.all(ds == value)
    assert np.all(sg == (1.0 if value > 0.5 else -1.0))


def test_sign_threshold_and_ghosts(oracle_lib):
    yo, xo, zo = 18, 30, 6
    py, pz, sep = 15, 20, 4
    px = sep + 4 * zo
    tmpa = voxels((yo, xo, zo), 3)
    sg, ds = O.ls_frontend(tmpa, (py, px, pz), sep, 3.0, n_ac=5)
    assert np.array_equal(sg, np.where(ds > 0.5, 1.0, -1.0))
    # NoGrdBc3D state: every ghost equals its clamped interior cell
    assert np.array_equal(O.nogrdbc3d(ds), ds)


def test_rejects_a_tunnel_loop_outside_tmpb(oracle_lib):
    # pz + (px-1)//zo beyond xo planes: the reference loop would leave TmpB
    tmpa = np.ones((6, 4, 3))
    with pytest.raises(ValueError):
        O.ls_frontend(tmpa, (5, 2 + 4 * 3, 4), 2, 1.0, n_ac=0)
