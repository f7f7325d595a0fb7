"""Host side of the learned CG warm start (libelsim_hip.so, elsim_nn_*).

The network (3x3x3 convolutional encoder-decoder, bfloat16 weights and
activations, FP32 accumulation on MFMA) runs and trains on the GPU through
the C-ABI; this module holds the parameter layout, checkpoints and a small
training loop over device-resident samples.
"""
import ctypes as C
import json
import os

import numpy as np

from ._lib import DeviceArray, check, lib

LAYERS = (  # name, cin, cout -- same order as the C-ABI flat parameter vector
    ("e0", 16, 16), ("e1", 16, 32), ("e2", 32, 64), ("e3", 64, 64),
    ("d2", 128, 64), ("d1", 96, 32), ("d0", 48, 16),
)
IN_CH = 16
FEATURES = ("psi_p", "residual", "depth", "log10_kappa", "dphi_prev")  # channels 0..4 of the input


def to_bf16(x):
    """float array -> bfloat16 bits (uint16), round to nearest even."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16(h):
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


def param_count():
    return int(lib().elsim_nn_param_count())


def flatten(p):
    """dict (name.w (cout, 27, cin), name.b, out.w (16,), out.b (1,)) -> flat FP32."""
    parts = []
    for name, cin, cout in LAYERS:
        parts += [np.asarray(p[name + ".w"], np.float32).reshape(cout, 27, cin).ravel(),
                  np.asarray(p[name + ".b"], np.float32).reshape(cout)]
    parts += [np.asarray(p["out.w"], np.float32).reshape(16), np.asarray(p["out.b"], np.float32).reshape(1)]
    v = np.concatenate(parts)
    assert v.size == param_count()
    return v


def unflatten(v):
    p, o = {}, 0
    for name, cin, cout in LAYERS:
        n = 27 * cin * cout
        p[name + ".w"] = v[o:o + n].reshape(cout, 27, cin)
        o += n
        p[name + ".b"] = v[o:o + cout]
        o += cout
    p["out.w"] = v[o:o + 16]
    p["out.b"] = v[o + 16:o + 17]
    return p


def he_init(seed):
    """He-uniform convolution weights, zero biases, small output layer."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, cin, cout in LAYERS:
        lim = np.sqrt(6.0 / (27 * cin))
        p[name + ".w"] = rng.uniform(-lim, lim, (cout, 27, cin))
        p[name + ".b"] = np.zeros(cout)
    p["out.w"] = rng.uniform(-0.1, 0.1, 16)
    p["out.b"] = np.zeros(1)
    return p


class WarmStartNet:
    """The network's device state (parameters, gradients, Adam moments)."""

    def __init__(self, params=None, seed=0):
        self.h = C.c_int64(0)
        check(lib().elsim_nn_create(C.byref(self.h)), "elsim_nn_create")
        self.meta = {}
        self.set_params(params if params is not None else flatten(he_init(seed)))

    def set_params(self, v):
        if isinstance(v, dict):
            v = flatten(v)
        v = np.ascontiguousarray(v, np.float32)
        check(lib().elsim_nn_set_params(C.byref(self.h), v.ctypes.data), "elsim_nn_set_params")

    def params(self):
        v = np.empty(param_count(), np.float32)
        check(lib().elsim_nn_get_params(C.byref(self.h), v.ctypes.data), "elsim_nn_get_params")
        return v

    def grads(self):
        v = np.empty(param_count(), np.float32)
        check(lib().elsim_nn_get_grads(C.byref(self.h), v.ctypes.data), "elsim_nn_get_grads")
        return v

    def forward(self, dims, x, pred, in_used=IN_CH):
        """x: device bf16 [N][in_used] (in_used 16, or 8 when input channels
        8..15 are zero: the first layer then runs on 8-channel slabs); pred:
        device FP32 [N]."""
        if in_used == IN_CH:
            check(lib().elsim_nn_forward(C.byref(self.h), *dims, x.ptr, pred.ptr), "elsim_nn_forward")
        else:
            check(lib().elsim_nn_forward_used(C.byref(self.h), *dims, x.ptr, in_used, pred.ptr),
                  "elsim_nn_forward_used")

    def loss_grad(self, dims, x, target, pred):
        loss = C.c_double()
        check(lib().elsim_nn_loss_grad(C.byref(self.h), *dims, x.ptr, target.ptr, pred.ptr, C.byref(loss)),
              "elsim_nn_loss_grad")
        return loss.value

    def train_forward(self, dims, x, pred):
        check(lib().elsim_nn_train_forward(C.byref(self.h), *dims, x.ptr, pred.ptr), "elsim_nn_train_forward")

    def energy_loss(self, pred, target, kap):
        """Squared A-norm error (phP operator of the ghosted FP32 coefficient
        kap) of pred against target; leaves its gradient for backward()."""
        loss = C.c_double()
        check(lib().elsim_nn_energy_loss(C.byref(self.h), pred.ptr, target.ptr, kap.ptr, C.byref(loss)),
              "elsim_nn_energy_loss")
        return loss.value

    def backward(self, x):
        check(lib().elsim_nn_backward(C.byref(self.h), x.ptr), "elsim_nn_backward")

    def adam(self, lr):
        check(lib().elsim_nn_adam(C.byref(self.h), C.c_float(lr)), "elsim_nn_adam")

    def save(self, path, **meta):
        m = dict(self.meta, **meta)
        np.savez_compressed(path, params=self.params(), meta=np.array(json.dumps(m)))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            net = cls(params=z["params"])
            net.meta = json.loads(str(z["meta"]))
        return net

    def close(self):
        if self.h.value:
            lib().elsim_nn_destroy(C.byref(self.h))
            self.h = C.c_int64(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warmstart.npz")
