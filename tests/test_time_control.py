"""The SBM program's time-loop control (SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3
.f90:358-383): the library's host function against the oracle over a grid of
inputs, and both pinned by values the reference fixes -- the REAL(4) products
for dt, the tier factors, the step the loop starts with (sCrnt = trgI,
SBM:351), and the REAL(4) thresholds, at inputs that fall between a
threshold's REAL(4) and REAL(8) values (a control written with REAL(8)
literals picks the other tier there)."""
import itertools

import numpy as np
import pytest

import elsim_amd as E
import oracle as O

DT_EARLY = 9.902343153953552e-4  # REAL(4) 1.05625e-02*0.125*0.75
DT_LATE = 6.601562490686774e-4  # REAL(4) 1.05625e-02*0.125*0.5
VSR0 = E.reference_params().vsr


def both(*a):
    g, o = E.time_control(*a), O.time_control(*a)
    assert g == o, (a, g, o)
    return g


def test_dt_products():
    f = np.float32
    assert DT_EARLY == float(f(f(f(1.05625e-02) * f(0.125)) * f(0.75)))
    assert DT_LATE == float(f(f(f(1.05625e-02) * f(0.125)) * f(0.5)))
    assert E.reference_params().dt == DT_EARLY


def test_first_step_runs_at_the_base_rate(oracle_lib):
    # sCrnt = trgI before the first step: the smallest Vsr tier (x0.5), and at
    # Xrf below 0.12 the x2 of the first dt tier -> exactly Vsr0
    for trgi in (-3.5e-2, 1.0, 7.25e3):
        for xrf in (0.0, 0.05, 0.1):
            assert both(xrf, trgi, trgi, VSR0) == (DT_EARLY, VSR0)


@pytest.mark.parametrize("xrf,dt,fx", [(0.0, DT_EARLY, 2), (0.1199999990, DT_LATE, 10), (0.12, DT_LATE, 10),
                                       (0.1199999, DT_EARLY, 2), (0.2, DT_LATE, 10), (0.2000000030, DT_LATE, 15),
                                       (0.5, DT_LATE, 15), (0.971, DT_LATE, 15)])
def test_lithiation_tiers(oracle_lib, xrf, dt, fx):
    # REAL(4) 0.12 = 0.119999997..., 0.20 = 0.200000003...
    trgi = 2.0
    assert both(xrf, trgi, trgi, 1.0) == (dt, 0.5 * fx)


@pytest.mark.parametrize("dc,fv", [(0.0, 0.5), (0.069, 0.5), (0.07, 0.5), (0.0700000004, 1.0), (0.19, 1.0),
                                   (0.2, 1.0), (0.2000000040, 3.0), (0.39, 3.0), (0.4, 3.0), (0.4000000061, 5.0),
                                   (0.74, 5.0), (0.75, 8.0), (3.0, 8.0)])
def test_current_tiers(oracle_lib, dc, fv):
    # |trgI| = 1: REAL(4) 0.07 = 0.0700000003, 0.20 = 0.200000003, 0.40 =
    # 0.400000006, 0.75 exact; the comparisons are strict
    for trgi, sgn in ((1.0, 1.0), (1.0, -1.0), (-1.0, 1.0)):
        assert both(0.5, trgi - sgn * dc, trgi, 1.0) == (DT_LATE, fv * 15)


def test_monotone_and_scale_free(oracle_lib):
    """Vsr grows with the current error and scales with Vsr0; dt depends on
    Xrf only; the tiers depend on |trgI - sCrnt| / |trgI| only (checked at
    power-of-two scalings, which change no rounding)."""
    xs = [0.0, 0.05, 0.119, 0.13, 0.199, 0.21, 0.6, 0.97]
    rs = np.linspace(0, 1.2, 49)
    for xrf in xs:
        v = [both(xrf, 1.0 - r, 1.0, VSR0) for r in rs]
        assert len({d for d, _ in v}) == 1
        assert all(a[1] <= b[1] for a, b in zip(v, v[1:]))
        for r, (d, vs) in zip(rs, v):
            assert both(xrf, 8.0 * (1.0 - r), 8.0, 0.25 * VSR0) == (d, 0.25 * vs)
            assert both(xrf, -(1.0 - r) / 4, -0.25, VSR0) == (d, vs)


def test_grid_against_oracle(oracle_lib):
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(0, 1, 40), [0.12, 0.2, np.nextafter(np.float32(0.12), 0, dtype=np.float32)]])
    for xrf, trgi in itertools.product(xs, (-2.5e-3, 1.0, 4.0e4)):
        for sc in rng.uniform(-1, 3, 6) * trgi:
            both(float(xrf), float(sc), trgi, VSR0)


def test_rejects_null():
    import ctypes as C
    d = C.c_double(0.0)
    assert E.lib().elsim_sbm_time_control(None, C.byref(d), C.byref(d), C.byref(d), C.byref(d), C.byref(d)) < 0
