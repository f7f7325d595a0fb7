"""Pins of the oracle's level-set family (LevelSet_3D.f90:164-322,
RM_UCNCT_3D.f90:180-337): bit-equality with the reference's own compiled
subroutines (golden fixtures, and oracle/_ref where built), plus closed-form
behaviour (distance to a plane, saturation, sign preservation)."""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DT = {"ls": 0.1, "rm": float(np.float32(0.1))}


@pytest.mark.parametrize("which", ["ls", "rm"])
def test_golden_fixtures_bit_exact(oracle_lib, which):
    z = np.load(os.path.join(GOLD, f"reinit_{which}.npz"))
    assert np.array_equal(O.redist3d(int(z["lit"]), z["sg"], np.zeros_like(z["sg"]), DT[which]), z["f_out"])
    assert np.array_equal(O.absgrd3duw(z["sg_rand"], z["f_rand"]), z["grad_out"])
    assert np.array_equal(O.nogrdbc3d(z["f_rand"]), z["nogrd_out"])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref (reference build) not present")
@pytest.mark.parametrize("which", ["ls", "rm"])
@pytest.mark.parametrize("seed", [0, 1])
def test_vs_reference_build(oracle_lib, which, seed):
    rng = np.random.default_rng(seed)
    shape = tuple(rng.integers(4, 14, 3))
    g = tuple(s + 2 for s in shape)
    sg = O.nogrdbc3d(np.where(rng.standard_normal(g) > 0.3, 1.0, -1.0))
    f0 = np.zeros(g)
    lit = int(rng.integers(5, 30))
    assert np.array_equal(O.redist3d(lit, sg, f0, DT[which]), O.ref_redist3d(lit, sg, f0, which))
    f = rng.standard_normal(g)
    sgr = np.sign(rng.standard_normal(g))
    assert np.array_equal(O.absgrd3duw(sgr, f), O.ref_absgrd3duw(sgr, f, which))


def test_plane_distance(oracle_lib):
    """A planar interface reinitialises to the exact signed distance."""
    shape = (6, 40, 5)
    g = tuple(s + 2 for s in shape)
    j = np.arange(g[1])[None, :, None] * np.ones(g)
    sg = np.where(j > 20, 1.0, -1.0)
    lit = 120
    f = O.redist3d(lit, sg, np.zeros(g), 0.1)
    d = j - 20.5
    core = (np.abs(d) < 6)
    assert np.abs(f - d)[core].max() < 0.02
    # saturation: |f| never exceeds 0.1 (lit - 1) + 1e-8 (the first sweep only
    # moves every cell to +-1e-8)
    assert np.abs(f).max() <= 0.1 * (lit - 1) + 1e-8 + 1e-12


def test_sign_preserved(oracle_lib):
    rng = np.random.default_rng(3)
    g = (12, 10, 9)
    sg = O.nogrdbc3d(np.where(rng.standard_normal(g) > 0, 1.0, -1.0))
    f = O.redist3d(15, sg, np.zeros(g), 0.1)
    assert np.all(np.sign(f) == sg)
