"""Host-side checks of the C-ABI library (no GPU calls): it loads and exports
every entry point include/elsim.h declares, the parameter struct layout, the
reference constants, the synthetic generators and the record format."""
import ctypes as C
import os

import numpy as np
import pytest

import elsim_amd as E
import oracle as O
from elsim_amd import fileio


def test_exports_every_declared_symbol():
    L = E.lib()
    names = E.exported_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_param_layout():
    # 33 doubles + 10 int32, no padding
    assert C.sizeof(E.SbmParams) == 33 * 8 + 10 * 4
    assert C.sizeof(E.StepStats) == 6 * 4 + 8 * 8 + 2 * 4 + 4 * 8


def test_reference_constants():
    """REAL(4)/REAL(8) literal kinds resolved as the reference computes them
    (SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3.f90:88-136)."""
    p = E.reference_params()
    assert p.dt == 9.902343153953552e-4
    assert p.rho == 0.031199999153614044
    assert p.frd == 96485.3359375
    assert p.bvp == -0.10000000149011612
    assert p.bve == -0.4494999945163727
    assert p.kap_scale == 3.299999952316284
    assert p.dab_c0 == -7.019999980926514
    assert p.dab_c2 == 0.05000000074505806
    assert p.kfw_cref == 0.0010000000474974513
    assert p.cst1 == pytest.approx(38.68173185437648, rel=1e-15)
    assert p.d0 == pytest.approx(0.004886868470654905, rel=1e-15)
    assert p.tc1 == pytest.approx(1.44375, rel=1e-15)
    assert p.tc2 == pytest.approx(106.61652342362515, rel=1e-15)


def test_sphere_pack_deterministic_and_connected(oracle_lib):
    a, na, fa = E.sphere_pack(40, 7, 4.0, 8.0)
    b, nb, fb = E.sphere_pack(40, 7, 4.0, 8.0)
    c, _, _ = E.sphere_pack(40, 8, 4.0, 8.0)
    assert np.array_equal(a, b) and na == nb and fa == fb
    assert not np.array_equal(a, c)
    assert 0.35 <= fa <= 0.6
    # every solid voxel reaches the current collector (high d2 face) and every
    # pore voxel the separator (low d2 face), through 6-connected paths
    import scipy.ndimage as ndi
    solid = a > 0
    lab, _ = ndi.label(solid)
    assert set(np.unique(lab[solid])) <= set(np.unique(lab[:, -1, :])) - {0}
    lab, _ = ndi.label(~solid)
    assert set(np.unique(lab[~solid])) <= set(np.unique(lab[:, 0, :])) - {0}
    # sdf is the distance to the nearest sphere surface (|grad| ~ 1 near it)
    g = np.gradient(a.astype(np.float64))
    gn = np.sqrt(sum(x ** 2 for x in g))
    band = np.abs(a) < 2
    assert np.median(gn[band]) == pytest.approx(1.0, abs=0.1)


def test_ellipsoids_and_tunnel():
    e, ne, fe = E.ellipsoid_pack(48, 3, 4.0, 10.0, cone_deg=15.0, target=0.45)
    assert ne > 5 and 0.2 <= fe <= 0.5
    t = E.carve_tunnel(e, 11, radius=3.0)
    assert np.all(t <= e)
    # the tunnel opens a pore path through the d2 extent
    import scipy.ndimage as ndi
    lab, _ = ndi.label(t <= 0)
    assert set(np.unique(lab[:, 0, :])) & set(np.unique(lab[:, -1, :])) - {0}


def test_record_roundtrip(tmp_path):
    a = np.random.default_rng(0).standard_normal((5, 4, 3)).astype(np.float32)
    p = tmp_path / "x.dat"
    fileio.write_record(p, a)
    raw = open(p, "rb").read()
    assert len(raw) == 4 + a.nbytes + 4
    assert int.from_bytes(raw[:4], "little") == a.nbytes
    assert raw[4:-4] == np.asfortranarray(a).tobytes(order="F")
    assert np.array_equal(fileio.read_record(p, a.shape, np.float32), a)
    assert fileio.output_name("CP1000", 0, 1)[:6] == "CP1000"
    assert fileio.output_name("PP", 3, 12) == "PP10031012.dat"
