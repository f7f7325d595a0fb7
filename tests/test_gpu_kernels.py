"""GPU parity of the individual kernels against the FP64 CPU oracle, through
the C-ABI.  Integer/byte work and FP64 arithmetic without transcendentals are
compared bit for bit; FP32 variants against their own FP32 restatement."""
import ctypes as C

import numpy as np
import pytest

import elsim_amd as E
from elsim_amd._lib import iref, ptr

pytestmark = pytest.mark.gpu


def rand_fields(shape, seed):
    rng = np.random.default_rng(seed)
    g = tuple(s + 2 for s in shape)
    p = rng.standard_normal(g)
    kap = rng.uniform(1e-7, 3.0, g)
    pse = rng.uniform(1e-6, 1.0, g)
    return p, kap, pse


# shapes cover partial waves, wave and tile edges in d1 (64, 128, 256 < 300),
# partial row blocks in d2 and partial plane blocks in d3
SHAPES = [(7, 5, 6), (64, 9, 17), (70, 33, 20), (128, 6, 18), (300, 7, 5)]


@pytest.mark.parametrize("op", [E.OP_CNE, E.OP_PHP, E.OP_PHE])
@pytest.mark.parametrize("shape", SHAPES)
def test_matvec_f64_bit_exact(gpu, oracle_lib, op, shape):
    O = oracle_lib
    p, kap, pse = rand_fields(shape, 1 + op)
    dt, dx = 1e-3, 0.7
    scale = dt / (dx * dx) if op == E.OP_CNE else 1.0 / (dx * dx)
    ref = O.apply_op(op, kap, pse, O.prj_bc(op, p), dt, dx)
    dp, dk, ds = (E.DeviceArray.from_host(O.flat(a)) for a in (p, kap, pse))
    out = E.DeviceArray(int(np.prod(shape)), np.float64)
    dot = E.DeviceArray(1, np.float64)
    E.check(gpu.elsim_dev_matvec_f64(op, *shape, dp.ptr, dk.ptr, ds.ptr, scale, out.ptr, dot.ptr))
    got = O.unflat(out.to_host(), shape)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    pin = O.prj_bc(op, p)[1:-1, 1:-1, 1:-1]
    assert dot.to_host()[0] == pytest.approx(float(np.sum(pin * ref)), rel=1e-12)


@pytest.mark.parametrize("op", [E.OP_CNE, E.OP_PHP, E.OP_PHE])
@pytest.mark.parametrize("shape", SHAPES)
def test_matvec_stored_ghosts_bit_exact(gpu, oracle_lib, op, shape):
    """The operator of the initial residual reads every ghost of p as stored."""
    O = oracle_lib
    p, kap, pse = rand_fields(shape, 11 + op)
    dt, dx = 1e-3, 0.7
    scale = dt / (dx * dx) if op == E.OP_CNE else 1.0 / (dx * dx)
    ref = O.apply_op(op, kap, pse, p, dt, dx)
    dp, dk, ds = (E.DeviceArray.from_host(O.flat(a)) for a in (p, kap, pse))
    out = E.DeviceArray(int(np.prod(shape)), np.float64)
    E.check(gpu.elsim_dev_matvec_stored_f64(op, *shape, dp.ptr, dk.ptr, ds.ptr, scale, out.ptr))
    got = O.unflat(out.to_host(), shape)
    assert np.array_equal(got, ref), np.abs(got - ref).max()


@pytest.mark.parametrize("op", [E.OP_CNE, E.OP_PHP, E.OP_PHE])
def test_matvec_f32(gpu, oracle_lib, op):
    O = oracle_lib
    shape = (40, 12, 9)
    p, kap, pse = rand_fields(shape, 7)
    dt, dx = 1e-3, 1.0
    scale = dt / (dx * dx) if op == E.OP_CNE else 1.0 / (dx * dx)
    ref = O.apply_op(op, kap, pse, O.prj_bc(op, p), dt, dx)
    f32 = lambda a: O.flat(a, np.float32)
    dp, dk, ds = (E.DeviceArray.from_host(f32(a)) for a in (p, kap, pse))
    out = E.DeviceArray(int(np.prod(shape)), np.float32)
    E.check(gpu.elsim_dev_matvec_f32(op, *shape, dp.ptr, dk.ptr, ds.ptr, C.c_float(scale), out.ptr, None))
    got = O.unflat(out.to_host(), shape).astype(np.float64)
    # FP32 stencil of O(1) data: error ~ a few ulp of the largest term
    assert np.abs(got - ref).max() <= 64 * np.finfo(np.float32).eps * np.abs(kap).max() * np.abs(p).max()


def apply_op_f32(op, kap, pse, pg, scale):
    """The 7-point operator in IEEE single precision, term by term in the
    order of the GPU kernels (SBM:515-521, 670-676, 751-757): face
    coefficients 0.5 (k_nb + k_c), then +f2 dp - f1 dp + f4 dp - f3 dp + f6 dp
    - f5 dp.  pg carries the homogeneous ghost rule (oracle prj_bc)."""
    h = np.float32(0.5)
    n = tuple(s - 2 for s in pg.shape)

    def sh(a, ax, o):
        sl = [slice(1, 1 + n[0]), slice(1, 1 + n[1]), slice(1, 1 + n[2])]
        sl[ax] = slice(1 + o, 1 + o + n[ax])
        return a[tuple(sl)]

    x, k0 = sh(pg, 0, 0), sh(kap, 0, 0)
    f = lambda ax, o: (sh(kap, ax, o) + k0) * h  # noqa: E731
    sv = f(0, 1) * (sh(pg, 0, 1) - x) - f(0, -1) * (x - sh(pg, 0, -1))
    sv = sv + f(1, 1) * (sh(pg, 1, 1) - x)
    sv = sv - f(1, -1) * (x - sh(pg, 1, -1))
    sv = sv + f(2, 1) * (sh(pg, 2, 1) - x)
    sv = sv - f(2, -1) * (x - sh(pg, 2, -1))
    return sh(pse, 0, 0) * x - scale * sv if op == E.OP_CNE else scale * sv


# d1 lengths: not a multiple of 4 (the one-voxel-a-lane operator), and
# multiples of 4 for the four-voxels-a-lane one: 64- and 128-wide segments,
# one full 256 tile, a partial second tile (300: the d1 tile-edge loads) and
# two full tiles (512)
SHAPES_F32 = [(7, 5, 6), (70, 33, 20), (64, 9, 17), (100, 7, 5), (128, 6, 18), (256, 20, 19), (300, 7, 5),
              (512, 5, 9)]


@pytest.mark.parametrize("op", [E.OP_CNE, E.OP_PHP, E.OP_PHE, 3])
@pytest.mark.parametrize("shape", SHAPES_F32)
def test_matvec_f32_bit_exact(gpu, oracle_lib, op, shape):
    O = oracle_lib
    p, kap, pse = rand_fields(shape, 21 + op)
    scale = np.float32(1e-3 / 0.49 if op == E.OP_CNE else 1.0 / 0.49)
    pg = O.prj_bc(op, p.astype(np.float32).astype(np.float64)).astype(np.float32)
    k32, s32 = kap.astype(np.float32), pse.astype(np.float32)
    ref = apply_op_f32(op, k32, s32, pg, scale)
    dp, dk, ds = (E.DeviceArray.from_host(O.flat(a, np.float32)) for a in (p, kap, pse))
    out = E.DeviceArray(int(np.prod(shape)), np.float32)
    dot = E.DeviceArray(1, np.float64)
    E.check(gpu.elsim_dev_matvec_f32(op, *shape, dp.ptr, dk.ptr, ds.ptr, C.c_float(scale), out.ptr, dot.ptr))
    got = O.unflat(out.to_host(), shape)
    assert np.array_equal(got, ref), (np.abs(got - ref).max(), np.argwhere(got != ref)[:5])
    want = float(np.sum(pg[1:-1, 1:-1, 1:-1].astype(np.float64) * ref.astype(np.float64)))
    assert dot.to_host()[0] == pytest.approx(want, rel=1e-12)


def sign_field(shape, seed):
    rng = np.random.default_rng(seed)
    g = tuple(s + 2 for s in shape)
    x = np.indices(g).astype(float)
    c = np.array(g, float) / 2 + rng.uniform(-1, 1, 3)
    r = np.sqrt(((x - c[:, None, None, None]) ** 2).sum(0))
    return np.where(r < min(g) / 3, -1.0, 1.0)


@pytest.mark.parametrize("shape", [(12, 10, 9), (67, 13, 11)])
def test_redist_f64_bit_exact(gpu, oracle_lib, shape):
    O = oracle_lib
    sg = O.nogrdbc3d(sign_field(shape, 3))
    f0 = np.zeros_like(sg)
    lit = 23
    for dt, in ((0.1,), (float(np.float32(0.1)),)):
        ref = O.redist3d(lit, sg, f0, dt)
        f = O.flat(f0)
        E.check(gpu.elsim_redist3d_xyz(iref(lit), iref(shape[0]), iref(shape[1]), iref(shape[2]), ptr(O.flat(sg)),
                                       ptr(f), C.byref(C.c_double(dt))))
        assert np.array_equal(O.unflat(f, sg.shape), ref)


def test_redist_f32_bit_exact(gpu, oracle_lib):
    O = oracle_lib
    shape = (33, 20, 14)
    sg = O.nogrdbc3d(sign_field(shape, 5))
    sg8 = sg.astype(np.int8)
    f0 = np.zeros(sg.shape, np.float32)
    ref = O.redist3d_f32(40, sg8, f0)
    ds = E.DeviceArray.from_host(O.flat(sg8, np.int8))
    df = E.DeviceArray.from_host(O.flat(f0, np.float32))
    E.check(gpu.elsim_dev_redist3d_f32(40, *shape, ds.ptr, df.ptr, C.c_float(np.float32(0.1))))
    E.check(gpu.elsim_synchronize())
    assert np.array_equal(O.unflat(df.to_host(), sg.shape), ref)


def test_absgrd_and_nogrdbc(gpu, oracle_lib):
    O = oracle_lib
    shape = (19, 8, 7)
    rng = np.random.default_rng(4)
    sg = np.sign(rng.standard_normal(tuple(s + 2 for s in shape)))
    sg[3, 3, 3] = 0.0
    f = rng.standard_normal(sg.shape)
    ref = O.absgrd3duw(sg, f)
    gr = np.zeros(int(np.prod(shape)))
    E.check(gpu.elsim_absgrd3duw_xyz(iref(shape[0]), iref(shape[1]), iref(shape[2]), ptr(O.flat(sg)), ptr(O.flat(f)),
                                     ptr(gr)))
    assert np.array_equal(O.unflat(gr, shape), ref)
    x = O.flat(f)
    E.check(gpu.elsim_nogrdbc3d(iref(shape[0]), iref(shape[1]), iref(shape[2]), ptr(x)))
    assert np.array_equal(O.unflat(x, f.shape), O.nogrdbc3d(f))


def test_table_bit_exact(gpu, oracle_lib):
    O = oracle_lib
    tab = E.tables(101)
    rng = np.random.default_rng(0)
    c = np.concatenate([rng.uniform(-0.1, 1.1, 4000), np.linspace(0, 1, 101), [0.0, 1.0, -0.0, 1e-6, 0.999999]])
    c = c[: (c.size // 6) * 6]
    shape = (3, 2, c.size // 6)
    for t in range(4):
        cord, vals = tab[2 * t * 101:(2 * t + 1) * 101], tab[(2 * t + 1) * 101:(2 * t + 2) * 101]
        fn = np.zeros(c.size)
        E.check(gpu.elsim_tbl_chk_pln_3da(iref(shape[0]), iref(shape[1]), iref(shape[2]), iref(101), ptr(cord),
                                          ptr(vals), ptr(np.ascontiguousarray(c)), ptr(fn)))
        ref = np.array([O.table(x, cord, vals) for x in c])
        assert np.array_equal(fn, ref)


def ad_case(shape, seed, frac=0.55):
    sdf, _, _ = E.sphere_pack(shape, seed, 3.0, 6.0, target=frac, hardcore=0.9, connect=False, dcut=12.0)
    return sdf.astype(np.float64)


@pytest.mark.parametrize("phase", [0, 1])
@pytest.mark.parametrize("shape", [(24, 30, 20), (65, 17, 33), (260, 5, 17), (300, 40, 36)])
def test_accdiff_bit_exact(gpu, oracle_lib, phase, shape):
    """Odd sizes: a d1 extent past one wave and past one 256-wide tile, an odd
    d2 extent (a half row tile) and d3 extents that end inside a plane tile;
    the last case has tiles in every direction that start, stop and resume
    changing while the front crosses them (the skipping of unchanged tiles)."""
    O = oracle_lib
    rd = ad_case(shape, 11)
    term = 151
    ref = O.accdiff_phase(rd, phase, term)
    cn = np.zeros(int(np.prod(shape)))
    E.check(gpu.elsim_accdiff_phase(iref(shape[0]), iref(shape[1]), iref(shape[2]), ptr(O.flat(rd)), iref(phase),
                                    iref(term), ptr(cn)))
    assert np.array_equal(O.unflat(cn, shape), ref)


@pytest.mark.parametrize("phase", [0, 1])
@pytest.mark.parametrize("shape", [(24, 30, 20), (65, 17, 33)])
def test_labels_bit_exact(gpu, oracle_lib, phase, shape):
    O = oracle_lib
    rd = ad_case(shape, 5 + shape[0])
    lab_ref, con_ref = O.label_connected(rd, phase)
    lab = np.zeros(int(np.prod(shape)), np.int32)
    con = np.zeros_like(lab)
    E.check(gpu.elsim_label_connected(iref(shape[0]), iref(shape[1]), iref(shape[2]), ptr(O.flat(rd)), iref(phase),
                                      ptr(lab), ptr(con)))
    assert np.array_equal(O.unflat(lab, shape), lab_ref)
    assert np.array_equal(O.unflat(con, shape), con_ref)


def test_labels_match_converged_accdiff(gpu, oracle_lib):
    """Once the AD front has covered every reachable voxel, the threshold
    Cn >= 1e-4 (RM_UCNCT_3D.f90:80) equals the labeller's connected set."""
    O = oracle_lib
    shape = (20, 24, 16)
    rd = ad_case(shape, 21)
    for phase in (0, 1):
        cn = np.zeros(int(np.prod(shape)))
        E.check(gpu.elsim_accdiff_phase(iref(shape[0]), iref(shape[1]), iref(shape[2]), ptr(O.flat(rd)), iref(phase),
                                        iref(1501), ptr(cn)))
        mask = O.unflat(cn, shape) >= np.float32(1e-4)
        con = np.zeros(int(np.prod(shape)), np.int32)
        lab = np.zeros_like(con)
        E.check(gpu.elsim_label_connected(iref(shape[0]), iref(shape[1]), iref(shape[2]), ptr(O.flat(rd)),
                                          iref(phase), ptr(lab), ptr(con)))
        assert np.array_equal(mask, O.unflat(con, shape).astype(bool))


def test_rm_threshold(gpu):
    rng = np.random.default_rng(2)
    shape = (9, 8, 7)
    cn = rng.uniform(0, 2e-4, shape)
    psi = rng.uniform(0, 1, shape)
    out = np.zeros(int(np.prod(shape)))
    E.check(gpu.elsim_rm_threshold(iref(9), iref(8), iref(7), ptr(np.ascontiguousarray(cn.ravel(order="F"))),
                                   ptr(np.ascontiguousarray(psi.ravel(order="F"))), ptr(out)))
    ref = np.where(cn < np.float32(1e-4), 0.0, psi)
    assert np.array_equal(out.reshape(shape, order="F"), ref)
