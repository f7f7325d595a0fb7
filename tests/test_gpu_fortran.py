"""The Fortran host path end to end on the GPU, against the oracle:
sbm_run (iso_c_binding) reads the reference's record formats, runs the SBM
time loop (fixed or adaptive dt) and writes OutPut3D_A-style FP32 records and
the TmFx file; geom_run runs the geometry chain LevelSet_3D -> AccDiff_3D ->
RM_UCNCT_3D -> 2x refinement through the drop-in subroutines."""
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
            if isinstance(v, bool):
                v = ".TRUE." if v else ".FALSE."
            elif isinstance(v, str):
                v = repr(v)
            f.write(f"  {k} = {v},\n")
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
    # the volume fractions, once before the loop (SBM:347-348)
    for ff, fid in (("PS", 0), ("PL", 1)):
        got = fileio.read_record(tmp_path / fileio.output_name(ff, 0, 1), (n, n, n), np.float32)
        assert np.allclose(got, o.get(fid)[1:-1, 1:-1, 1:-1], rtol=1e-6, atol=1e-7), ff


def _r4(x):
    return float(np.float32(x))


def oracle_time_loop(o, prm, nsteps, cr=6.0):
    """The reference's time loop (SBM:341-877) over the oracle's step, with the
    oracle's time-loop control: the TmFx lines and the frames it writes."""
    f = np.float32
    sc = o.scalars()
    xrf, bve = sc["xrf"], sc["bve"]
    trgi = _r4(0.95) * sc["tpp"] * prm.rho / (3600 / cr)
    vsr0, scrnt, cnt, k, tm = prm.vsr, trgi, 1, 0, 0.0
    lines, frames, stop = [], {}, None

    def frame(fr):
        frames[fr] = {fid: o.get(fid)[1:-1, 1:-1, 1:-1] for fid in (F_CNP, F_CNE, F_PHP, F_PHE)}

    while xrf < _r4(0.971):
        if k >= nsteps:
            lines.append((cnt, k, tm, xrf, scrnt, trgi, bve))
            frame(cnt)
            stop = "cap"
            break
        k += 1
        prm.dt, prm.vsr = O.time_control(xrf, scrnt, trgi, vsr0)
        o.set_params(prm)
        st = o.step()
        sc = o.scalars()
        xrf, bve, scrnt = sc["xrf"], sc["bve"], st.scrnt
        tm += prm.dt
        if xrf > float(f(f(0.02) + f(f(cnt - 1) * f(0.005)))):
            if cnt % 10 == 1:
                frame(cnt)
            lines.append((cnt, k, tm, xrf, scrnt, trgi, bve))
            cnt += 1
        if bve >= prm.bvp + _r4(0.005) or xrf > _r4(0.97):
            lines.append((cnt, k, tm, xrf, scrnt, trgi, bve))
            frame(cnt)
            stop = "test"
            break
    return lines, frames, stop


@pytest.mark.parametrize("case", ["cadence_and_cap", "stop_test"])
def test_sbm_run_adaptive_matches_oracle(tmp_path, case):
    """sbm_run with adaptive = T against the same loop over the oracle: the
    TmFx lines (frame, iteration, time, Xrf, current, target, BvE) and every
    frame's fields."""
    assert os.path.exists(EXE), "Fortran driver not built"
    n = 20
    sdf, _, _ = E.sphere_pack(n, 3, 4.0, 8.0)
    prm = E.synthetic_params(vsr=0.02)
    tab = E.tables()
    if case == "cadence_and_cap":
        cnp, nsteps = E.initial_cnp(n, 2), 3
    else:  # starts just above the stop test's Xrf > 0.97
        cnp, nsteps = E.initial_cnp(n, 2, mean=0.9703, amp=1e-4), 5
    fileio.write_record(tmp_path / "ds.dat", sdf, np.float32)
    fileio.write_record(tmp_path / "dl.dat", -sdf, np.float32)
    fileio.write_record(tmp_path / "tab.dat", tab.reshape((101, 8), order="F"), np.float64)
    fileio.write_record(tmp_path / "cnp.dat", cnp, np.float64)
    namelist(tmp_path / "run.nml", dict(n1=n, n2=n, n3=n, nsteps=nsteps, mode=E.MODE_MIXED,
                                        sdf_s=str(tmp_path / "ds.dat"), sdf_l=str(tmp_path / "dl.dat"),
                                        cnp_file=str(tmp_path / "cnp.dat"), tables_file=str(tmp_path / "tab.dat"),
                                        outdir=str(tmp_path), outpf="GBQ105", adaptive=True,
                                        asmb=case == "stop_test"), prm)
    r = subprocess.run([EXE, str(tmp_path / "run.nml")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    o = O.OracleSbm(sdf, -sdf, E.synthetic_params(vsr=0.02), tab)
    o.set(F_CNP, cnp)
    lines, frames, stop = oracle_time_loop(o, o.params, nsteps)
    assert stop == ("cap" if case == "cadence_and_cap" else "test")
    # list-directed records of 7 items (the compiler may wrap a record over lines)
    tok = open(tmp_path / "GBQ105TmFx.txt").read().split()
    assert len(tok) % 7 == 0, tok
    got = [tok[i:i + 7] for i in range(0, len(tok), 7)]
    assert len(got) == len(lines), (got, lines)
    for g, w in zip(got, lines):
        assert (int(g[0]), int(g[1])) == w[:2]
        v = [float(x) for x in g[2:]]
        assert v[0] == pytest.approx(w[2], rel=1e-14)  # time: a sum of the control's dt values
        assert v[3] == pytest.approx(w[5], rel=1e-12)  # target current from tPP
        assert abs(v[1] - w[3]) <= 1e-8, (g, w)  # Xrf
        assert abs(v[2] - w[4]) <= 1e-6 * max(abs(w[4]), abs(w[5])), (g, w)  # current, against the target's scale
        assert abs(v[4] - w[6]) <= 1e-12, (g, w)  # BvE moves by +-Vsr dt per step
    for fr, fields in frames.items():
        for ff, fid in (("CP", F_CNP), ("CE", F_CNE), ("PP", F_PHP), ("PE", F_PHE)):
            name = f"{ff}ASMB{1000 + fr}.dat" if case == "stop_test" else fileio.output_name(ff, 0, fr)
            got_f = fileio.read_record(tmp_path / name, (n, n, n), np.float32)
            ref = fields[fid]
            e = np.linalg.norm(got_f - ref) / np.linalg.norm(ref)
            assert e <= 1e-6, (fr, ff, e)
    names = sorted(os.listdir(tmp_path))
    rk = "ASMB" if case == "stop_test" else "1000"
    assert f"PS{rk}1001.dat" in names and f"PL{rk}1001.dat" in names, names
    if rk == "ASMB":
        assert not any(x.startswith(("CP1000", "PS1000")) for x in names), names
    o.close()


def test_geom_run_matches_oracle(tmp_path):
    """geom_run (Fortran: LS_FrontEnd, ReDist3D_XYZ, AccDiff_Phase, RM_Phase,
    ReDist3D_XYZ_R4, Refine2_SDF) writes every record of the geometry chain
    bit-exact to the oracle's chain, under the reference's file names."""
    from elsim_amd import frontend as F
    from elsim_amd import pipeline as P
    exe = os.path.join(os.path.dirname(E.__file__), "lib", "geom_run")
    assert os.path.exists(exe), "Fortran geometry driver not built"
    image, (py, px, pz), sep, tr, fr, term = (20, 30, 8), (16, 4 + 4 * 8, 20), 4, 4.5, 5, 301
    sdf, _, _ = E.sphere_pack(image, 5, 2.0, 4.0, target=0.5, hardcore=0.85, connect=False)
    tmpa = (sdf > 0).astype(np.float64)
    np.savetxt(tmp_path / "img.txt", tmpa.ravel(order="F"), fmt="%.1f")
    with open(tmp_path / "g.nml", "w") as f:
        f.write(f"&geom image='{tmp_path / 'img.txt'}', yo={image[0]}, xo={image[1]}, zo={image[2]}, py={py}, "
                f"px={px}, pz={pz}, sep={sep}, nac=50, fr={fr}, tr={tr!r}, lit=175, term={term}, "
                f"outdir='{tmp_path}' /\n")
    r = subprocess.run([exe, str(tmp_path / "g.nml")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    shape, g, frm = (py, px, pz), f"{py}x{px}x{pz}", f"{100 + fr}"
    sg, _ = O.ls_frontend(tmpa, shape, sep, tr, F.REF_NAC)
    l24 = O.redist3d(175, sg, np.zeros_like(sg), 0.1)[1:-1, 1:-1, 1:-1]
    rd = lambda name, dt=np.float64, s=shape: fileio.read_record(tmp_path / name, s, dt)  # noqa: E731
    assert np.array_equal(rd(f"L24Z{1000 + fr}.dat"), l24)
    for ph, pre in ((0, "SOD"), (1, "LIQ")):
        cn = O.accdiff_phase(l24, ph, term)
        assert np.array_equal(rd(f"{pre}_3D_CN_{g}_Z_II_{frm}.dat"), cn), pre
        ls = O.redist3d(175, O.rm_phase(l24, cn, ph), np.zeros_like(sg), P.RM_DT)[1:-1, 1:-1, 1:-1]
        lsname, fine = ("GRP_LS", "DS") if ph == 0 else ("POR_LS", "DL")
        assert np.array_equal(rd(f"{lsname}_{g}_Z_II_{frm}.dat"), ls), lsname
        got = rd(f"GrMic_II_{fine}_{2 * py}x{2 * px}x{2 * pz}_B{frm}.dat", np.float32, (2 * py, 2 * px, 2 * pz))
        assert np.array_equal(got, O.refine2(ls)), fine


@pytest.mark.parametrize("phase", [0, 1])
def test_dropins_from_fortran(tmp_path, phase):
    """NoGrdBc3D, AbsGrd3DUW_XYZ, Tbl_Chk_Pln_3DA, Label_Connected and
    RM_Threshold called from Fortran (tests/fortran/dropin_check.f90) with the
    reference's argument lists: bit-exact against the oracle."""
    exe = os.path.join(os.path.dirname(E.__file__), "lib", "dropin_check")
    assert os.path.exists(exe), "Fortran drop-in check program not built"
    py, px, pz, sz = 13, 9, 7, 101
    rng = np.random.default_rng(40 + phase)
    g = (py + 2, px + 2, pz + 2)
    f = rng.standard_normal(g)
    sg = rng.choice([-1.0, 0.0, 1.0], size=g, p=[0.45, 0.1, 0.45])
    c = rng.uniform(-0.2, 1.2, (py, px, pz))
    c[0, 0, 0] = float(np.float32(1e-4))
    psa = rng.uniform(0, 1, (py, px, pz))
    rdin, _, _ = E.sphere_pack((py, px, pz), 9 + phase, 1.5, 3.0, target=0.5, hardcore=0.85, connect=False, dcut=8.0)
    rdin = rdin.astype(np.float64)
    cord = np.linspace(0.0, 1.0, sz)
    vals = rng.standard_normal(sz)
    for name, a in (("f", f), ("sg", sg), ("c", c), ("psa", psa), ("rd", rdin), ("cord", cord), ("vals", vals)):
        fileio.write_record(tmp_path / f"{name}.dat", a, np.float64)
    r = subprocess.run([exe, str(tmp_path), str(py), str(px), str(pz), str(sz), str(phase)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rd = lambda name, s, dt=np.float64: fileio.read_record(tmp_path / name, s, dt)  # noqa: E731
    n = (py, px, pz)
    assert np.array_equal(rd("gr_out.dat", n), O.absgrd3duw(sg, f))
    assert np.array_equal(rd("f_out.dat", g), O.nogrdbc3d(f))
    fn = np.array([O.table(x, cord, vals) for x in c.ravel(order="F")]).reshape(n, order="F")
    assert np.array_equal(rd("fn_out.dat", n), fn)
    assert np.array_equal(rd("psb_out.dat", n), np.where(c < np.float32(1e-4), 0.0, psa))
    lab, con = O.label_connected(rdin, phase)
    assert np.array_equal(rd("lab_out.dat", n, np.int32), lab)
    assert np.array_equal(rd("con_out.dat", n, np.int32), con)


def test_rm_dropin_library(tmp_path):
    """RM_UCNCT_3D's call of ReDist3D_XYZ (RM_UCNCT_3D.f90:100, 148) linked
    unchanged against libelsim_rm.so (tests/fortran/rm_check.f90): the REAL(4)
    pseudo-time step of RM_UCNCT_3D.f90:236, bit-exact against the oracle's
    restatement and, where it was built, the reference's own subroutine."""
    exe = os.path.join(os.path.dirname(E.__file__), "lib", "rm_check")
    assert os.path.exists(exe), "RM drop-in check program not built"
    py, px, pz, lit = 11, 14, 9, 15
    rng = np.random.default_rng(77)
    g = (py + 2, px + 2, pz + 2)
    sg = O.nogrdbc3d(np.where(rng.uniform(size=g) > 0.5, 1.0, -1.0))
    f = np.zeros(g)
    for name, a in (("sg", sg), ("f", f)):
        fileio.write_record(tmp_path / f"{name}.dat", a, np.float64)
    r = subprocess.run([exe, str(tmp_path), str(lit), str(py), str(px), str(pz)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = fileio.read_record(tmp_path / "f_out.dat", g, np.float64)
    want = O.redist3d(lit, sg, f, float(np.float32(0.1)))
    assert np.array_equal(got, want)
    assert not np.array_equal(got, O.redist3d(lit, sg, f, 0.1))  # the LS step gives other values
    if O.ref_available():
        assert np.array_equal(got, O.ref_redist3d(lit, sg, f, "rm"))
