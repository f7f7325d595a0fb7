"""The rank-output stitcher sbm_asmb (SBMES_3D_Cmbn_MPI_v2.f90:3-137), host
file I/O only: per-rank FP32 records cut from a global array by numpy's
array_split -- which splits as MyDmnCmpn_YZ (SBM:947-982) does, the first
MOD(g, parts) parts one cell longer -- are stitched back bit-exactly into
<FF>ASMB<1000+frame>.dat."""
import os
import subprocess

import numpy as np
import pytest

import elsim_amd as E
from elsim_amd import fileio

EXE = os.path.join(os.path.dirname(E.__file__), "lib", "sbm_asmb")
pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="sbm_asmb not built (needs flang)")


def rank_files(d, ff, frame, glob, rnb, cnb):
    """Cut glob (gy, gx, gz) into the rnb x cnb rank boxes: rank = (ddC-1) rnb
    + (ddR-1), d3 part ddR, d1 part ddC; write <ff><1000+rank><1000+frame>.dat."""
    ys = np.array_split(np.arange(glob.shape[0]), cnb)
    zs = np.array_split(np.arange(glob.shape[2]), rnb)
    for c, y in enumerate(ys):
        for r, z in enumerate(zs):
            rank = c * rnb + r
            part = glob[y[0]:y[-1] + 1, :, z[0]:z[-1] + 1]
            fileio.write_record(d / f"{ff}{1000 + rank}{1000 + frame}.dat", part, np.float32)


def run(d, **kw):
    with open(d / "a.nml", "w") as f:
        f.write("&asmb\n")
        for k, v in kw.items():
            if isinstance(v, (list, tuple)):
                v = ",".join(repr(x) for x in v)
            elif isinstance(v, str):
                v = repr(v)
            f.write(f"  {k} = {v},\n")
        f.write("/\n")
    return subprocess.run([EXE, str(d / "a.nml")], capture_output=True, text=True, timeout=120)


def test_stitches_fields_and_frames(tmp_path):
    gy, gx, gz, rnb, cnb = 13, 5, 11, 3, 4
    rng = np.random.default_rng(1)
    want = {}
    for ff in ("CP", "PE"):
        for fr in (3, 5, 7):
            g = rng.standard_normal((gy, gx, gz)).astype(np.float32)
            want[ff, fr] = g
            rank_files(tmp_path, ff, fr, g, rnb, cnb)
    r = run(tmp_path, gy=gy, gx=gx, gz=gz, rnb=rnb, cnb=cnb, fields=["CP", "PE"], fb=3, fe=7, itv=2,
            indir=str(tmp_path), outdir=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    for (ff, fr), g in want.items():
        p = tmp_path / f"{ff}ASMB{1000 + fr}.dat"
        raw = open(p, "rb").read()
        assert len(raw) == 8 + g.nbytes
        assert np.array_equal(fileio.read_record(p, g.shape, np.float32), g)


def test_reference_decomposition(tmp_path):
    """The reference's 186 x . x 320 grid over 12 x 16 ranks (d2 cut to 2 cells):
    d1 parts 10 x 12 + 6 x 11 cells, d3 parts 8 x 27 + 4 x 26 cells."""
    gy, gx, gz, rnb, cnb = 186, 2, 320, 12, 16
    assert [len(a) for a in np.array_split(np.arange(gy), cnb)] == [12] * 10 + [11] * 6
    assert [len(a) for a in np.array_split(np.arange(gz), rnb)] == [27] * 8 + [26] * 4
    g = np.arange(gy * gx * gz, dtype=np.float32).reshape((gy, gx, gz), order="F")
    rank_files(tmp_path, "CE", 135, g, rnb, cnb)
    r = run(tmp_path, gy=gy, gx=gx, gz=gz, rnb=rnb, cnb=cnb, fields=["CE"], fb=135, fe=135, itv=10,
            indir=str(tmp_path), outdir=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert np.array_equal(fileio.read_record(tmp_path / "CEASMB1135.dat", g.shape, np.float32), g)


def test_missing_rank_file_fails(tmp_path):
    g = np.zeros((4, 3, 4), np.float32)
    rank_files(tmp_path, "CP", 1, g, 2, 2)
    os.remove(tmp_path / "CP10031001.dat")
    r = run(tmp_path, gy=4, gx=3, gz=4, rnb=2, cnb=2, fields=["CP"], indir=str(tmp_path), outdir=str(tmp_path))
    assert r.returncode != 0
    assert not (tmp_path / "CPASMB1001.dat").exists()


def test_rejects_more_ranks_than_cells(tmp_path):
    r = run(tmp_path, gy=3, gx=3, gz=4, rnb=2, cnb=4, indir=str(tmp_path), outdir=str(tmp_path))
    assert r.returncode != 0
