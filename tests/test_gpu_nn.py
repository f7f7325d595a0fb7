"""GPU parity of the warm-start network's kernels against the CPU oracle
(oracle/nn_oracle.py, FP64 with bfloat16 rounding at the GPU's storage
points).  Differences come only from FP32 accumulation order: a stored
bfloat16 value may differ by one bfloat16 ulp, FP32 results by ~1e-6."""
import ctypes as C

import numpy as np
import pytest

from elsim_amd import nn as N
from elsim_amd._lib import DeviceArray, check
from oracle import nn_oracle as NO

pytestmark = pytest.mark.gpu

BF_ULP = 2.0 ** -7


def dev(a):
    return DeviceArray.from_host(np.ascontiguousarray(a))


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def cl(x):  # (n3, n2, n1, C) channels-last -> flat [voxel][C]
    return np.ascontiguousarray(x).reshape(-1, x.shape[-1])


@pytest.mark.parametrize("cin,cout,shape", [(16, 16, (20, 12, 9)), (32, 64, (16, 8, 8)), (48, 16, (17, 5, 6)),
                                            (96, 32, (8, 8, 4)), (128, 64, (16, 4, 3)), (8, 16, (20, 12, 9)),
                                            (8, 32, (9, 7, 5))])
def test_conv_forward(gpu, cin, cout, shape):
    n1, n2, n3 = shape
    rng = np.random.default_rng(cin + cout)
    x = NO.bf16(rng.standard_normal((n3, n2, n1, cin)))
    w = NO.bf16(rng.uniform(-1, 1, (cout, 27, cin)) / np.sqrt(27 * cin))
    b = rng.uniform(-0.1, 0.1, cout).astype(np.float32)
    ref = NO.bf16(np.maximum(NO.conv3(x, w, b.astype(np.float64)), 0))
    out = DeviceArray((n1 * n2 * n3 * cout,), np.uint16)
    xd, wd, bd = dev(N.to_bf16(cl(x))), dev(N.to_bf16(w.reshape(cout, -1))), dev(b)  # kept alive over the call
    check(gpu.elsim_nn_conv3(n1, n2, n3, xd.ptr, cin, wd.ptr, bd.ptr, cout, 1, out.ptr))
    got = N.from_bf16(out.to_host()).reshape(n3, n2, n1, cout).astype(np.float64)
    assert np.all(np.abs(got - ref) <= BF_ULP * np.abs(ref) + 1e-6)
    assert np.mean(got == ref) > 0.95


@pytest.mark.parametrize("cin,cout,shape", [(48, 16, (16, 8, 8)), (128, 64, (8, 4, 4)), (16, 32, (9, 7, 5))])
def test_conv_backward_data(gpu, cin, cout, shape):
    n1, n2, n3 = shape
    rng = np.random.default_rng(7 + cin)
    dz = NO.bf16(rng.standard_normal((n3, n2, n1, cout)))
    w = NO.bf16(rng.uniform(-1, 1, (cout, 27, cin)))
    ref = NO.conv3_bwd_data(dz, w)
    wb = np.ascontiguousarray(w[:, ::-1, :].transpose(2, 1, 0)).reshape(cin, 27 * cout)  # [ci][tap' cout + co]
    dx = DeviceArray((n1 * n2 * n3 * cin,), np.float32)
    zd, wd = dev(N.to_bf16(cl(dz))), dev(N.to_bf16(wb))
    check(gpu.elsim_nn_conv3_bwd_data(n1, n2, n3, zd.ptr, cout, wd.ptr, cin, dx.ptr))
    got = dx.to_host().reshape(n3, n2, n1, cin)
    assert rel_l2(got, ref) < 2e-6


@pytest.mark.parametrize("cin,cout,shape", [(48, 16, (16, 16, 16)), (128, 64, (8, 8, 8)), (16, 32, (24, 16, 12))])
def test_conv_backward_weight(gpu, cin, cout, shape):
    n1, n2, n3 = shape
    rng = np.random.default_rng(11 + cin)
    x = NO.bf16(rng.standard_normal((n3, n2, n1, cin)))
    dz = NO.bf16(rng.standard_normal((n3, n2, n1, cout)))
    ref = NO.conv3_bwd_weight(x, dz)
    dw = DeviceArray((cout * 27 * cin,), np.float32)
    xd, zd = dev(N.to_bf16(cl(x))), dev(N.to_bf16(cl(dz)))
    check(gpu.elsim_nn_conv3_bwd_weight(n1, n2, n3, xd.ptr, cin, zd.ptr, cout, dw.ptr))
    got = dw.to_host().reshape(cout, 27, cin)
    assert rel_l2(got, ref) < 1e-5


def network_case(shape, seed):
    n1, n2, n3 = shape
    rng = np.random.default_rng(seed)
    p = NO.init_params(seed)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.uniform(0.0, 0.1, p[k].shape)
    x = NO.bf16(rng.standard_normal((n3, n2, n1, 16)))
    x[..., 5:] = 0.0
    target = rng.standard_normal((n3, n2, n1)) * 0.1
    return p, x, target


@pytest.mark.parametrize("shape", [(16, 16, 16), (24, 16, 8)])
def test_network_forward_and_gradients(gpu, shape):
    n1, n2, n3 = shape
    p, x, target = network_case(shape, 3)
    pred_ref, cache = NO.forward(p, x)
    g_ref = NO.backward(p, cache, pred_ref, target)
    net = N.WarmStartNet(params=N.flatten(p))
    xd = dev(N.to_bf16(cl(x)))
    td = dev(target.astype(np.float32).ravel())
    pd = DeviceArray((n1 * n2 * n3,), np.float32)
    loss = net.loss_grad((n1, n2, n3), xd, td, pd)
    pred = pd.to_host().reshape(n3, n2, n1)
    assert rel_l2(pred, pred_ref) < 2e-2
    assert loss == pytest.approx(NO.loss(pred_ref, target), rel=2e-2)
    g = N.unflatten(net.grads())
    for k in g_ref:
        assert rel_l2(g[k], g_ref[k]) < 5e-2, k
    # The inference forward: the same prediction up to its last decoder
    # layer, which convolves the upsampled half at half resolution with the
    # weights folded per voxel parity (sums of bf16 weights rounded to bf16)
    # and sums in another order.
    pd2 = DeviceArray((n1 * n2 * n3,), np.float32)
    net.forward((n1, n2, n3), xd, pd2)
    e16 = rel_l2(pd2.to_host(), pd.to_host())
    assert rel_l2(pd2.to_host().reshape(n3, n2, n1), pred_ref) < 2e-2
    # channels 5..15 are zero: the forward on rows of channels 0..7 (first
    # layer on 8-channel slabs) differs by FP32 summation order alone
    x8 = dev(N.to_bf16(cl(x[..., :8])))
    net.forward((n1, n2, n3), x8, pd2, in_used=8)
    e8 = rel_l2(pd2.to_host(), pd.to_host())
    print("inference forward vs training forward: rel l2 %.2e (16 channels), %.2e (8)" % (e16, e8))
    assert e16 < 5e-3 and e8 < 5e-3
    assert rel_l2(pd2.to_host().reshape(n3, n2, n1), pred_ref) < 2e-2
    net.close()


def test_adam_step_matches_oracle(gpu):
    shape = (16, 16, 8)
    n1, n2, n3 = shape
    p, x, target = network_case(shape, 5)
    net = N.WarmStartNet(params=N.flatten(p))
    xd, td = dev(N.to_bf16(cl(x))), dev(target.astype(np.float32).ravel())
    pd = DeviceArray((n1 * n2 * n3,), np.float32)
    net.loss_grad(shape, xd, td, pd)
    g = N.unflatten(net.grads().astype(np.float64))
    p0 = N.unflatten(net.params().astype(np.float64))
    net.adam(1e-3)
    p1 = N.unflatten(net.params().astype(np.float64))
    pr = {k: v.copy() for k, v in p0.items()}
    m = {k: np.zeros_like(v) for k, v in p0.items()}
    v = {k: np.zeros_like(v) for k, v in p0.items()}
    NO.adam(pr, g, m, v, 1, 1e-3)
    for k in p1:
        assert np.allclose(p1[k], pr[k], rtol=1e-5, atol=1e-7), k
    net.close()


def test_training_reduces_loss(gpu):
    """A few Adam steps on one fixed sample lower the loss (the whole
    forward/backward/update chain is consistent)."""
    shape = (16, 16, 16)
    n1, n2, n3 = shape
    rng = np.random.default_rng(0)
    x = NO.bf16(rng.standard_normal((n3, n2, n1, 16)))
    x[..., 5:] = 0.0
    target = 0.5 * x[..., 0] - 0.25 * x[..., 1]  # a local function of the inputs
    net = N.WarmStartNet(seed=1)
    xd, td = dev(N.to_bf16(cl(x))), dev(target.astype(np.float32).ravel())
    pd = DeviceArray((n1 * n2 * n3,), np.float32)
    losses = []
    for _ in range(60):
        losses.append(net.loss_grad(shape, xd, td, pd))
        net.adam(3e-3)
    assert losses[-1] < 0.2 * losses[0], losses[::10]
    net.close()


def test_energy_loss_and_gradients(gpu, oracle_lib):
    """Squared A-norm loss e.(-A e)/N (A: homogeneous phP operator of kap,
    unit scale) and the gradients it sends through the network."""
    import elsim_amd as E
    shape = (16, 8, 8)
    n1, n2, n3 = shape
    p, x, target = network_case(shape, 9)
    rng = np.random.default_rng(9)
    kap = rng.uniform(1e-7, 1.0, (n1 + 2, n2 + 2, n3 + 2))  # ghosted, [i, j, k]
    net = N.WarmStartNet(params=N.flatten(p))
    xd, td = dev(N.to_bf16(cl(x))), dev(target.astype(np.float32).ravel())
    kd = dev(oracle_lib.flat(kap, np.float32))
    pd = DeviceArray((n1 * n2 * n3,), np.float32)
    net.train_forward(shape, xd, pd)
    loss = net.energy_loss(pd, td, kd)
    net.backward(xd)
    g = N.unflatten(net.grads())
    # oracle: the same forward, then dpred = -2 A e / N in FP64
    pred_ref, cache = NO.forward(p, x)
    pred_gpu = pd.to_host().reshape(n3, n2, n1).astype(np.float64)
    e = (pred_gpu - target.astype(np.float32).astype(np.float64)).transpose(2, 1, 0)  # -> [i, j, k]
    eg = np.zeros((n1 + 2, n2 + 2, n3 + 2))
    eg[1:-1, 1:-1, 1:-1] = e
    kap32 = kap.astype(np.float32).astype(np.float64)
    Ae = oracle_lib.apply_op(E.OP_PHP, kap32, None, oracle_lib.prj_bc(E.OP_PHP, eg), 0.0, 1.0)
    N0 = n1 * n2 * n3
    assert loss == pytest.approx(-np.sum(e * Ae) / N0, rel=1e-5)
    dpred = (-2.0 * Ae / N0).transpose(2, 1, 0)  # -> (n3, n2, n1)
    g_ref = NO.backward(p, cache, pred_ref, target, dpred=dpred)
    for k in g_ref:
        assert rel_l2(g[k], g_ref[k]) < 5e-2, k
    net.close()
