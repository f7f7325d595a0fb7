"""Pins of the warm-start network's CPU oracle (no GPU): the convolution
against a brute-force loop, the adjoint identities of every backward
operator, and the full backward pass against central finite differences of
the unrounded forward."""
import numpy as np
import pytest

from oracle import nn_oracle as NO


def brute_conv(x, w, b):
    n3, n2, n1, ci = x.shape
    co = w.shape[0]
    out = np.zeros((n3, n2, n1, co))
    for z in range(n3):
        for y in range(n2):
            for xx in range(n1):
                s = b.copy()
                for t in range(27):
                    dx, dy, dz = t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1
                    zz, yy, x2 = z + dz, y + dy, xx + dx
                    if 0 <= zz < n3 and 0 <= yy < n2 and 0 <= x2 < n1:
                        s = s + w[:, t, :] @ x[zz, yy, x2]
                out[z, y, xx] = s
    return out


def test_conv_matches_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 4, 5, 2))
    w = rng.standard_normal((3, 27, 2))
    b = rng.standard_normal(3)
    assert np.allclose(NO.conv3(x, w, b), brute_conv(x, w, b), rtol=1e-12, atol=1e-12)


def test_tap_order_is_d1_fastest():
    # a weight only on tap 1 + 3*1 + 9*1 + 1 = (dx=+1, dy=0, dz=0) picks the d1 neighbour
    x = np.zeros((3, 3, 3, 1))
    x[1, 1, 2, 0] = 1.0
    w = np.zeros((1, 27, 1))
    w[0, 14, 0] = 1.0  # dx=+1: 2 + 3*1 + 9*1
    assert NO.conv3(x, w)[1, 1, 1, 0] == 1.0


def test_backward_operators_are_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 6, 8, 3))
    w = rng.standard_normal((5, 27, 3))
    y = rng.standard_normal((4, 6, 8, 5))
    lhs = np.sum(NO.conv3(x, w) * y)
    assert np.isclose(lhs, np.sum(x * NO.conv3_bwd_data(y, w)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(w * NO.conv3_bwd_weight(x, y)), rtol=1e-12)
    c = rng.standard_normal((2, 3, 4, 3))
    assert np.isclose(np.sum(NO.pool(x) * c), np.sum(x * NO.pool_bwd(c)), rtol=1e-12)
    assert np.isclose(np.sum(NO.up(c) * x), np.sum(c * NO.up_bwd(x)), rtol=1e-12)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-40, 3.0e38])
    r = NO.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -6  # tie to even (upwards)
    assert r[3] == -2.5
    assert np.all(np.isfinite(r))


@pytest.mark.parametrize("name", ["e0.w", "e3.w", "d2.w", "d0.b", "out.w", "e1.b"])
def test_backward_matches_finite_differences(name):
    rng = np.random.default_rng(3)
    p = NO.init_params(5)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.uniform(0.05, 0.2, p[k].shape)  # keep ReLUs away from their kink
    x0 = rng.standard_normal((8, 8, 8, 16))
    target = rng.standard_normal((8, 8, 8))
    pred, cache = NO.forward(p, x0, exact=True)
    g = NO.backward(p, cache, pred, target, exact=True)
    flat = p[name].reshape(-1)
    idx = rng.choice(flat.size, size=min(4, flat.size), replace=False)
    for i in idx:
        h = 1e-6 * max(1.0, abs(flat[i]))
        old = flat[i]
        flat[i] = old + h
        lp = NO.loss(NO.forward(p, x0, exact=True)[0], target)
        flat[i] = old - h
        lm = NO.loss(NO.forward(p, x0, exact=True)[0], target)
        flat[i] = old
        fd = (lp - lm) / (2 * h)
        an = g[name].reshape(-1)[i]
        assert an == pytest.approx(fd, rel=1e-5, abs=1e-10), (name, i)
