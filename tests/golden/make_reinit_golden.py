"""Write the level-set golden fixtures from the reference's own subroutines.

Run in the build container (it needs oracle/_ref, compiled by oracle/Makefile
from the reference's LevelSet_3D.f90 and RM_UCNCT_3D.f90):
    python tests/golden/make_reinit_golden.py
Outputs tests/golden/reinit_{ls,rm}.npz: inputs and the reference's outputs of
NoGrdBc3D, AbsGrd3DUW_XYZ and ReDist3D_XYZ on small grids.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402


def sign_field(shape, seed):
    rng = np.random.default_rng(seed)
    g = tuple(s + 2 for s in shape)
    x = np.indices(g).astype(float)
    c = np.array(g, float) / 2 + rng.uniform(-1, 1, 3)
    r = np.sqrt(((x - c[:, None, None, None]) ** 2).sum(0))
    sg = np.where(r < min(g) / 3, -1.0, 1.0)
    sg[:, 1, :] = 1.0  # a planar strip as well
    return O.ref_nogrdbc3d(sg)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref not built (make -C oracle)")
    out = os.path.dirname(os.path.abspath(__file__))
    shape = (11, 9, 7)
    rng = np.random.default_rng(2024)
    for which in ("ls", "rm"):
        sg = sign_field(shape, 1 if which == "ls" else 2)
        f0 = np.zeros_like(sg)
        lit = 40
        f = O.ref_redist3d(lit, sg, f0, which)
        fr = rng.standard_normal(sg.shape)
        sgr = np.sign(rng.standard_normal(sg.shape))
        sgr[2, 3, 4] = 0.0
        gr = O.ref_absgrd3duw(sgr, fr, which)
        nb = O.ref_nogrdbc3d(fr, which)
        np.savez_compressed(os.path.join(out, f"reinit_{which}.npz"), shape=np.array(shape), lit=lit, sg=sg,
                            f_out=f, f_rand=fr, sg_rand=sgr, grad_out=gr, nogrd_out=nb)
        print("wrote", which)


if __name__ == "__main__":
    main()
