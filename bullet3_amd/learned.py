"""The learned warm start's network over the C ABI (b3hMlp in include/b3hip.h).

A 3-layer MLP in -> 128 -> 128 -> out (GELU, bf16 weights and activations,
fp32 accumulation) that runs and trains on the device.  Parameters cross the
boundary as one packed float32 array: w1[in][128] b1[128] w2[128][128] b2[128]
w3[128][out] b3[out].

The reference has no counterpart: its solver always starts from zero impulses
(b3ConvertConstraint4.h:73,107,132).
"""
import ctypes

import numpy as np

from . import capi
from .capi import call, ptr

HIDDEN = 128


def param_shapes(in_dim, out_dim):
    return [("w1", (in_dim, HIDDEN)), ("b1", (HIDDEN,)), ("w2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
            ("w3", (HIDDEN, out_dim)), ("b3", (out_dim,))]


def unpack(flat, in_dim, out_dim):
    out, o = {}, 0
    for name, shp in param_shapes(in_dim, out_dim):
        n = int(np.prod(shp))
        out[name] = flat[o:o + n].reshape(shp)
        o += n
    return out


def pack(params, in_dim, out_dim):
    return np.concatenate([np.asarray(params[n], np.float32).ravel() for n, _ in param_shapes(in_dim, out_dim)])


class DeviceArray:
    """float32 device array filled from / read into numpy (through the C ABI)."""

    def __init__(self, arr_or_shape):
        if isinstance(arr_or_shape, np.ndarray):
            a = np.ascontiguousarray(arr_or_shape, np.float32)
            self.shape = a.shape
            self.buf = capi.DeviceBuffer(max(a.nbytes, 4))
            capi.htod(self.buf.ptr, a)
        else:
            self.shape = tuple(arr_or_shape)
            self.buf = capi.DeviceBuffer(max(4 * int(np.prod(self.shape)), 4))

    @property
    def ptr(self):
        return self.buf.ptr

    def numpy(self, rows=None):
        """the array (or its first `rows` rows) on the host"""
        shape = self.shape if rows is None else (min(int(rows), self.shape[0]),) + tuple(self.shape[1:])
        n = int(np.prod(shape))
        return capi.dtoh(self.buf.ptr, 4 * n, np.float32).reshape(shape).copy()

    def free(self):
        self.buf.free()


class WarmStartNet:
    def __init__(self, in_dim=24, out_dim=3, max_train_rows=8192, seed=1, stream=None):
        self.in_dim, self.out_dim = int(in_dim), int(out_dim)
        self.h = ctypes.c_void_p()
        call("b3h_mlp_create", self.in_dim, self.out_dim, int(max_train_rows), ctypes.c_uint(seed), stream,
             ctypes.byref(self.h))
        n = ctypes.c_int()
        call("b3h_mlp_num_params", self.h, ctypes.byref(n))
        self.num_params = n.value

    # -- parameters
    def get_params(self):
        out = np.zeros(self.num_params, np.float32)
        call("b3h_mlp_get_params", self.h, ptr(out))
        return out

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        assert flat.size == self.num_params
        call("b3h_mlp_set_params", self.h, ptr(flat))

    def set_input_normalization(self, shift, scale):
        shift = np.ascontiguousarray(shift, np.float32)
        scale = np.ascontiguousarray(scale, np.float32)
        assert shift.size == self.in_dim and scale.size == self.in_dim
        call("b3h_mlp_set_input_normalization", self.h, ptr(shift), ptr(scale))

    def get_input_normalization(self):
        shift = np.zeros(self.in_dim, np.float32)
        scale = np.zeros(self.in_dim, np.float32)
        call("b3h_mlp_get_input_normalization", self.h, ptr(shift), ptr(scale))
        return shift, scale

    def save(self, path):
        shift, scale = self.get_input_normalization()
        np.savez(path, params=self.get_params(), in_dim=self.in_dim, out_dim=self.out_dim, shift=shift, scale=scale)

    def load(self, path):
        z = np.load(path)
        assert int(z["in_dim"]) == self.in_dim and int(z["out_dim"]) == self.out_dim
        self.set_params(z["params"])
        self.set_input_normalization(z["shift"], z["scale"])

    # -- compute
    def forward_dev(self, x_dev, rows, y_dev, ldx=None, ldy=None):
        call("b3h_mlp_forward", self.h, ctypes.c_void_p(x_dev), int(ldx or self.in_dim), int(rows),
             ctypes.c_void_p(y_dev), int(ldy or self.out_dim))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        dx = DeviceArray(x)
        dy = DeviceArray((x.shape[0], self.out_dim))
        self.forward_dev(dx.ptr, x.shape[0], dy.ptr, x.shape[1], self.out_dim)
        call("b3h_synchronize")
        y = dy.numpy()
        dx.free()
        dy.free()
        return y

    def loss_gradients(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        dx, dt = DeviceArray(x), DeviceArray(t)
        g = np.zeros(self.num_params, np.float32)
        loss = ctypes.c_float()
        call("b3h_mlp_loss_gradients", self.h, ctypes.c_void_p(dx.ptr), x.shape[1], ctypes.c_void_p(dt.ptr),
             t.shape[1], x.shape[0], ptr(g), ctypes.byref(loss))
        dx.free()
        dt.free()
        return loss.value, g

    def train_dev(self, x_dev, ldx, t_dev, ldt, rows, batch_rows, steps, lr, seed=0):
        hist = np.zeros(max(steps, 1), np.float32)
        call("b3h_mlp_train", self.h, ctypes.c_void_p(x_dev), int(ldx), ctypes.c_void_p(t_dev), int(ldt), int(rows),
             int(batch_rows), int(steps), ctypes.c_float(lr), ctypes.c_uint(seed), ptr(hist))
        return hist[:steps]

    def train(self, x, t, batch_rows, steps, lr, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        dx, dt = DeviceArray(x), DeviceArray(t)
        hist = self.train_dev(dx.ptr, x.shape[1], dt.ptr, t.shape[1], x.shape[0], batch_rows, steps, lr, seed)
        dx.free()
        dt.free()
        return hist

    def close(self):
        if self.h:
            call("b3h_mlp_destroy", self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
