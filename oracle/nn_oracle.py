"""CPU oracle of the warm-start network (test infrastructure only).

Only the tests, the smoke check and the benchmark's CPU leg may import this
module; the product path (elsim_amd) never does.

The network is this project's own (it has no counterpart in the reference):
a 4-level 3-D convolutional encoder-decoder, channels 16/32/64/64, 3x3x3
"same" convolutions (stride 1, zero padding), 2x2x2 average pooling down,
nearest-neighbour upsampling up, skip connections by channel concatenation
(upsampled channels first), ReLU after every convolution and a 1x1x1 output
convolution to one channel.  The GPU stores activations, convolution
weights and activation gradients in bfloat16 and accumulates in FP32; this
restatement computes in FP64 and rounds to bfloat16 at exactly the points
where the GPU stores a bfloat16 tensor (`rnd`), so its results differ from
the GPU's only by FP32 accumulation order.  With `exact=True` nothing is
rounded, which is what the finite-difference checks of the backward pass use.

Tensors are channels-last: an activation at level l is an array
(n3_l, n2_l, n1_l, C) (d1 fastest, as in the GPU's [voxel][channel] layout
with voxel = i + n1 (j + n2 k)).  Convolution weights are W[co, tap, ci] with
tap = dx + 3 (dy + 3 dz), offsets d* in {0, 1, 2} meaning -1, 0, +1 along
d1 (x), d2 (y), d3 (z).
"""
import numpy as np

LEVEL_CH = (16, 32, 64, 64)
IN_CH = 16  # input channels as stored (5 features, zero padded)
LAYERS = (  # name, cin, cout
    ("e0", 16, 16), ("e1", 16, 32), ("e2", 32, 64), ("e3", 64, 64),
    ("d2", 128, 64), ("d1", 96, 32), ("d0", 48, 16),
)


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


class Rounder:
    def __init__(self, exact):
        self.exact = exact

    def __call__(self, x):
        return np.asarray(x, np.float64) if self.exact else bf16(x)


def taps():
    """(dz, dy, dx) offsets in -1..1 for tap = dx+1 + 3 (dy+1 + 3 (dz+1))."""
    out = []
    for t in range(27):
        dx, dy, dz = t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1
        out.append((dz, dy, dx))
    return out


def shift(x, dz, dy, dx):
    """y[z, y, x] = x[z+dz, y+dy, x+dx] with zeros outside."""
    n3, n2, n1 = x.shape[:3]
    y = np.zeros_like(x)
    zs, ys, xs = slice(max(0, -dz), n3 - max(0, dz)), slice(max(0, -dy), n2 - max(0, dy)), slice(max(0, -dx),
                                                                                                 n1 - max(0, dx))
    zt, yt, xt = slice(max(0, dz), n3 - max(0, -dz)), slice(max(0, dy), n2 - max(0, -dy)), slice(max(0, dx),
                                                                                                 n1 - max(0, -dx))
    y[zs, ys, xs] = x[zt, yt, xt]
    return y


def conv3(x, w, b=None):
    """out[v, co] = b[co] + sum_{tap, ci} w[co, tap, ci] x[v + o_tap, ci] (FP64)."""
    out = np.zeros(x.shape[:3] + (w.shape[0],))
    for t, (dz, dy, dx) in enumerate(taps()):
        out += shift(x, dz, dy, dx) @ w[:, t, :].T
    if b is not None:
        out += b
    return out


def conv3_bwd_data(dz_, w):
    """Adjoint of conv3 in x: dx[u, ci] = sum_{tap, co} w[co, tap, ci] dz[u - o_tap, co]."""
    out = np.zeros(dz_.shape[:3] + (w.shape[2],))
    for t, (dz, dy, dx) in enumerate(taps()):
        out += shift(dz_, -dz, -dy, -dx) @ w[:, t, :]
    return out


def conv3_bwd_weight(x, dz_):
    """dw[co, tap, ci] = sum_v dz[v, co] x[v + o_tap, ci]."""
    co, ci = dz_.shape[3], x.shape[3]
    dw = np.zeros((co, 27, ci))
    dzf = dz_.reshape(-1, co)
    for t, (dz, dy, dx) in enumerate(taps()):
        dw[:, t, :] = dzf.T @ shift(x, dz, dy, dx).reshape(-1, ci)
    return dw


def pool(x):
    """2x2x2 average."""
    n3, n2, n1, c = x.shape
    return x.reshape(n3 // 2, 2, n2 // 2, 2, n1 // 2, 2, c).mean(axis=(1, 3, 5))


def pool_bwd(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_bwd(g):
    n3, n2, n1, c = g.shape
    return g.reshape(n3 // 2, 2, n2 // 2, 2, n1 // 2, 2, c).sum(axis=(1, 3, 5))


def init_params(seed=0):
    """He-uniform weights, zero biases (the GPU's initialisation is the same
    function of the seed only through this oracle in tests; the trainer
    draws its own)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, cin, cout in LAYERS:
        lim = np.sqrt(6.0 / (27 * cin))
        p[name + ".w"] = rng.uniform(-lim, lim, (cout, 27, cin))
        p[name + ".b"] = np.zeros(cout)
    p["out.w"] = rng.uniform(-0.25, 0.25, 16)
    p["out.b"] = np.zeros(1)
    return p


def forward(p, x0, exact=False):
    """Network forward.  x0: (n3, n2, n1, 16) input features.  Returns
    (pred (n3, n2, n1), cache)."""
    R = Rounder(exact)
    W = {k: (R(v) if k.endswith(".w") and not k.startswith("out") else np.asarray(v, np.float64))
         for k, v in p.items()}
    relu = lambda a: np.maximum(a, 0.0)
    c = {"x0": R(x0)}
    c["e0"] = R(relu(conv3(c["x0"], W["e0.w"], W["e0.b"])))
    c["p1"] = R(pool(c["e0"]))
    c["e1"] = R(relu(conv3(c["p1"], W["e1.w"], W["e1.b"])))
    c["p2"] = R(pool(c["e1"]))
    c["e2"] = R(relu(conv3(c["p2"], W["e2.w"], W["e2.b"])))
    c["p3"] = R(pool(c["e2"]))
    c["e3"] = R(relu(conv3(c["p3"], W["e3.w"], W["e3.b"])))
    c["cat2"] = np.concatenate([up(c["e3"]), c["e2"]], axis=3)
    c["d2"] = R(relu(conv3(c["cat2"], W["d2.w"], W["d2.b"])))
    c["cat1"] = np.concatenate([up(c["d2"]), c["e1"]], axis=3)
    c["d1"] = R(relu(conv3(c["cat1"], W["d1.w"], W["d1.b"])))
    c["cat0"] = np.concatenate([up(c["d1"]), c["e0"]], axis=3)
    c["d0"] = R(relu(conv3(c["cat0"], W["d0.w"], W["d0.b"])))
    pred = c["d0"] @ W["out.w"] + W["out.b"][0]
    c["W"] = W
    return pred, c


def loss(pred, target):
    return float(np.mean((pred - target) ** 2))


def backward(p, cache, pred, target, exact=False, dpred=None):
    """Gradients of loss(pred, target) = mean((pred - target)^2) (or of any
    loss whose gradient with respect to pred is `dpred`) with respect to
    every parameter, following the GPU's rounding points: the gradient
    reaching a convolution output is FP32 (here FP64), multiplied by the ReLU
    mask and rounded to bfloat16 before it enters the backward
    convolutions."""
    R = Rounder(exact)
    W, c = cache["W"], cache
    g = {}
    n = pred.size
    if dpred is None:
        dpred = 2.0 * (pred - target) / n
    g["out.w"] = np.tensordot(dpred, c["d0"], axes=([0, 1, 2], [0, 1, 2]))
    g["out.b"] = np.array([dpred.sum()])
    d_d0 = dpred[..., None] * W["out.w"]

    def layer(name, x_in, act, dact):
        dz = R(dact * (act > 0))
        g[name + ".w"] = conv3_bwd_weight(x_in, dz)
        g[name + ".b"] = dz.reshape(-1, dz.shape[3]).sum(0)
        return conv3_bwd_data(dz, W[name + ".w"])

    dcat0 = layer("d0", c["cat0"], c["d0"], d_d0)
    d_d1 = up_bwd(dcat0[..., :32])
    d_e0 = dcat0[..., 32:].copy()
    dcat1 = layer("d1", c["cat1"], c["d1"], d_d1)
    d_d2 = up_bwd(dcat1[..., :64])
    d_e1 = dcat1[..., 64:].copy()
    dcat2 = layer("d2", c["cat2"], c["d2"], d_d2)
    d_e3 = up_bwd(dcat2[..., :64])
    d_e2 = dcat2[..., 64:].copy()
    dp3 = layer("e3", c["p3"], c["e3"], d_e3)
    d_e2 += pool_bwd(dp3)
    dp2 = layer("e2", c["p2"], c["e2"], d_e2)
    d_e1 += pool_bwd(dp2)
    dp1 = layer("e1", c["p1"], c["e1"], d_e1)
    d_e0 += pool_bwd(dp1)
    layer("e0", c["x0"], c["e0"], d_e0)
    return g


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected), in place."""
    for k in p:
        m[k] = b1 * m[k] + (1 - b1) * g[k]
        v[k] = b2 * v[k] + (1 - b2) * g[k] ** 2
        mh = m[k] / (1 - b1 ** t)
        vh = v[k] / (1 - b2 ** t)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)
