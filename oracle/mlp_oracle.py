"""ORACLE -- test infrastructure only.

CPU restatement (numpy) of the learned warm start's network: a 3-layer MLP
in -> 128 -> 128 -> out, GELU in its tanh form, mean squared error, Adam.  The
reference has no such network (its solver cold-starts,
b3ConvertConstraint4.h:73,107,132); this file is the checker of the device
kernels in bullet3_amd/csrc/k_mlp.hip.

Two modes:
  * exact=True: float64 throughout, no rounding -- the textbook network.  Its
    analytic gradient is pinned by central finite differences
    (tests/test_mlp_oracle.py), GELU by its closed-form values and the erf form.
  * exact=False: the device's rounding points -- input, weights, both hidden
    activations and the back-propagated gradients that feed a matrix product
    are rounded to bf16 (nearest even); sums are float64 here, fp32 on the
    device (the tests allow for that difference).

Parameters use the packed layout of the C ABI: w1[in][128] b1[128]
w2[128][128] b2[128] w3[128][out] b3[out].
"""
import numpy as np

HIDDEN = 128
GELU_C = 0.7978845608028654
GELU_A = 0.044715


def bf16(x):
    """round float32 values to the nearest bfloat16 (ties to even), as float32"""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(GELU_C * (x + GELU_A * x ** 3)))


def gelu_grad(x):
    th = np.tanh(GELU_C * (x + GELU_A * x ** 3))
    return 0.5 * (1.0 + th) + 0.5 * x * (1.0 - th * th) * GELU_C * (1.0 + 3.0 * GELU_A * x * x)


def unpack(flat, in_dim, out_dim):
    shapes = [("w1", (in_dim, HIDDEN)), ("b1", (HIDDEN,)), ("w2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
              ("w3", (HIDDEN, out_dim)), ("b3", (out_dim,))]
    out, o = {}, 0
    for name, shp in shapes:
        n = int(np.prod(shp))
        out[name] = np.asarray(flat[o:o + n], np.float64).reshape(shp)
        o += n
    assert o == len(flat)
    return out


def pack(p):
    return np.concatenate([np.ravel(p[k]) for k in ("w1", "b1", "w2", "b2", "w3", "b3")])


def _r(x, exact):
    return np.asarray(x, np.float64) if exact else bf16(np.asarray(x, np.float32)).astype(np.float64)


def _f32(x, exact):
    return np.asarray(x, np.float64) if exact else np.asarray(x, np.float32).astype(np.float64)


def normalize(x, shift=None, scale=None, exact=False):
    """the network's input normalisation (x - shift) * scale (float32 on the device)"""
    if shift is None:
        return x
    if exact:
        return (np.asarray(x, np.float64) - shift) * scale
    return ((np.asarray(x, np.float32) - np.float32(shift)) * np.float32(scale)).astype(np.float32)


def forward(flat, x, in_dim, out_dim, exact=False, keep=False, shift=None, scale=None):
    p = unpack(flat, in_dim, out_dim)
    xb = _r(normalize(x, shift, scale, exact), exact)
    w1, w2, w3 = _r(p["w1"], exact), _r(p["w2"], exact), _r(p["w3"], exact)
    z1 = _f32(xb @ w1 + p["b1"], exact)
    h1 = _r(gelu(z1), exact)
    z2 = _f32(h1 @ w2 + p["b2"], exact)
    h2 = _r(gelu(z2), exact)
    y = _f32(h2 @ w3 + p["b3"], exact)
    if keep:
        return y, dict(xb=xb, w1=w1, w2=w2, w3=w3, z1=z1, h1=h1, z2=z2, h2=h2)
    return y


def loss_gradients(flat, x, t, in_dim, out_dim, exact=False, shift=None, scale=None):
    """mean squared error over rows x out and its gradient (packed layout)"""
    y, c = forward(flat, x, in_dim, out_dim, exact, keep=True, shift=shift, scale=scale)
    m = x.shape[0]
    d = y - np.asarray(t, np.float64)
    loss = float(np.mean(d * d))
    dy = _f32(2.0 * d / (m * out_dim), exact)
    dyb = _r(dy, exact)
    g = {}
    g["w3"] = c["h2"].T @ dyb
    g["b3"] = dy.sum(0)
    dz2 = _f32((dyb @ c["w3"].T) * gelu_grad(c["z2"]), exact)
    g2 = _r(dz2, exact)
    g["w2"] = c["h1"].T @ g2
    g["b2"] = dz2.sum(0)
    dz1 = _f32((g2 @ c["w2"].T) * gelu_grad(c["z1"]), exact)
    g1 = _r(dz1, exact)
    g["w1"] = c["xb"].T @ g1
    g["b1"] = dz1.sum(0)
    return loss, pack(g)


def adam(flat, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """one Adam step (step counts from 1); returns new (flat, m, v)"""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return flat - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
