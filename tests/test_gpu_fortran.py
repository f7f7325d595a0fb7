"""The Fortran host path end to end: sbm_run (Fortran, iso_c_binding) reads
the reference's record formats, steps on the GPU and writes OutPut3D_A-style
FP32 records, compared with the FP64 oracle; plus the drop-in ReDist3D_XYZ
called from Fortran."""
import os
import subprocess

import numpy as np
import pytest

import elsim_amd as E
import oracle as O
from elsim_amd import fileio
from elsim_amd._lib import F_CNE, F_CNP, F_PHE, F_PHP

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(E.__file__), "lib", "sbm_run")


def namelist(path, run, prm):
    with open(path, "w") as f:
        f.write("&run\n")
        for k, v in run.items():
            f.write(f"  {k} = {v!r},\n" if isinstance(v, str) else f"  {k} = {v},\n")
        f.write("/\n&params\n")
        for k, v in prm.as_dict().items():
            f.write(f"  prm%{k} = {v!r},\n")
        f.write("/\n")


def test_sbm_run_matches_oracle(tmp_path):
    assert os.path.exists(EXE), "Fortran driver not built"
    n, steps = 20, 2
    sdf, _, _ = E.sphere_pack(n, 3, 4.0, 8.0)
    prm = E.synthetic_params()
    tab = E.tables()
    cnp = E.initial_cnp(n, 2)
    fileio.write_record(tmp_path / "ds.dat", sdf, np.float32)
    fileio.write_record(tmp_path / "dl.dat", -sdf, np.float32)
    fileio.write_record(tmp_path / "tab.dat", tab.reshape((101, 8), order="F"), np.float64)
    fileio.write_record(tmp_path / "cnp.dat", cnp, np.float64)
    namelist(tmp_path / "run.nml", dict(n1=n, n2=n, n3=n, nsteps=steps, mode=E.MODE_MIXED,
                                        sdf_s=str(tmp_path / "ds.dat"), sdf_l=str(tmp_path / "dl.dat"),
                                        cnp_file=str(tmp_path / "cnp.dat"), tables_file=str(tmp_path / "tab.dat"),
                                        outdir=str(tmp_path), outpf="GBQ105"), prm)
    r = subprocess.run([EXE, str(tmp_path / "run.nml")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    o = O.OracleSbm(sdf, -sdf, prm, tab)
    o.set(F_CNP, cnp)
    for _ in range(steps):
        o.step()
    for ff, fid in (("CP", F_CNP), ("CE", F_CNE), ("PP", F_PHP), ("PE", F_PHE)):
        got = fileio.read_record(tmp_path / fileio.output_name(ff, 0, 1), (n, n, n), np.float32)
        ref = o.get(fid)[1:-1, 1:-1, 1:-1]
        e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert e <= 1e-6, (ff, e)
    line = open(tmp_path / "GBQ105TmFx.txt").read().split()
    assert len(line) == 7 and int(line[1]) == steps
