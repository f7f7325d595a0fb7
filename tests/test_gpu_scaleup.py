"""Level set and connectivity at 512^3 on an ellipsoid packing with a
random-walk tunnel: int32 component labels of both phases bit-exact to the
oracle's breadth-first labelling, the FP64 reinitialisation bit-exact to the
oracle's restatement of ReDist3D_XYZ (LevelSet_3D.f90:194-261) on a 96^3
sub-box, and the error of 200 FP32 sweeps against the FP64 ones at full
size."""
import ctypes as C

import numpy as np
import pytest

import elsim_amd as E

pytestmark = pytest.mark.gpu

N = 512
SWEEPS = 200


@pytest.fixture(scope="module")
def geometry():
    sdf, ns, frac = E.ellipsoid_pack(N, 3, 8.0, 24.0, cone_deg=15.0, target=0.5, overlap=0.3, dcut=12.0)
    sdf = E.carve_tunnel(sdf, 3, radius=3.0)
    print(f"ellipsoids {ns}, solid {frac:.3f}")
    return sdf


def ghosted(a):
    g = np.empty(tuple(s + 2 for s in a.shape), a.dtype)
    g[1:-1, 1:-1, 1:-1] = a
    for ax in range(3):  # mirror ghosts, d1 then d2 then d3 (corners included)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        src_lo = [slice(None)] * 3
        src_hi = [slice(None)] * 3
        lo[ax], src_lo[ax] = 0, 1
        hi[ax], src_hi[ax] = -1, -2
        g[tuple(lo)] = g[tuple(src_lo)]
        g[tuple(hi)] = g[tuple(src_hi)]
    return g


@pytest.mark.parametrize("phase", [0, 1])
def test_labels_512_bit_exact(gpu, oracle_lib, geometry, phase):
    rd = np.asfortranarray(geometry.astype(np.float64))
    ref_lab, ref_con = oracle_lib.label_connected(rd, phase)
    d_rd = E.DeviceArray.from_host(rd.ravel(order="F"))
    d_lab = E.DeviceArray((N ** 3,), np.int32)
    d_con = E.DeviceArray((N ** 3,), np.int32)
    E.check(gpu.elsim_dev_label_connected(N, N, N, d_rd.ptr, phase, d_lab.ptr, d_con.ptr))
    lab = d_lab.to_host().reshape((N, N, N), order="F")
    con = d_con.to_host().reshape((N, N, N), order="F")
    ncomp = len(np.unique(lab[lab >= 0]))
    print(f"phase {phase}: {ncomp} components, {con.mean():.4f} of the voxels face-connected")
    assert np.array_equal(lab, ref_lab)
    assert np.array_equal(con, ref_con)


def test_reinit_512_fp32_vs_fp64(gpu, oracle_lib, geometry):
    phi0 = np.tanh(geometry.astype(np.float64) / 2.0)  # a level set that is not a distance function
    f = ghosted(phi0)
    sg = np.sign(f)
    # FP64 on the GPU is the oracle's restatement bit for bit: check on a sub-box
    sub = tuple(slice(200, 296) for _ in range(3))
    fs = ghosted(phi0[sub])
    ref = oracle_lib.redist3d(SWEEPS, np.sign(fs), fs, 0.1)
    d_sg, d_f = E.DeviceArray.from_host(oracle_lib.flat(np.sign(fs))), E.DeviceArray.from_host(oracle_lib.flat(fs))
    E.check(gpu.elsim_dev_redist3d_f64(SWEEPS, 96, 96, 96, d_sg.ptr, d_f.ptr, C.c_double(0.1)))
    assert np.array_equal(oracle_lib.unflat(d_f.to_host(), fs.shape), ref)
    # full size: FP64 and FP32 (int8 sign) on the GPU
    d_sg = E.DeviceArray.from_host(oracle_lib.flat(sg))
    d_f = E.DeviceArray.from_host(oracle_lib.flat(f))
    E.check(gpu.elsim_dev_redist3d_f64(SWEEPS, N, N, N, d_sg.ptr, d_f.ptr, C.c_double(0.1)))
    f64 = d_f.to_host()
    del d_f, d_sg
    d_sg8 = E.DeviceArray.from_host(oracle_lib.flat(sg, np.int8))
    d_f32 = E.DeviceArray.from_host(oracle_lib.flat(f, np.float32))
    E.check(gpu.elsim_dev_redist3d_f32(SWEEPS, N, N, N, d_sg8.ptr, d_f32.ptr, C.c_float(0.1)))
    f32 = d_f32.to_host().astype(np.float64)
    err = np.abs(f32 - f64)
    band = np.abs(f64) <= 5.0
    eb = err[band]
    rms, mx, tail = float(np.sqrt(np.mean(eb ** 2))), float(eb.max()), float(np.mean(eb > 1e-3))
    print(f"FP32 vs FP64 after {SWEEPS} sweeps, band |phi|<=5 ({band.mean():.3f} of the box): rms {rms:.3e}, "
          f"max {mx:.3e}, fraction > 1e-3 {tail:.2e}; whole box max {err.max():.3e}")
    # FP32 carries ~7 digits, so the rms error stays at rounding level; the
    # upwind choices and the sign-flip guard are discrete, so a few voxels
    # take a different branch and move by up to a few hundredths of a voxel
    assert rms <= 1e-5
    assert mx <= 0.05
    assert tail <= 1e-5
    assert np.mean(np.sign(f32[band]) != np.sign(f64[band])) <= 1e-6
