"""Device network of the learned warm start (k_mlp.hip, through the C ABI)
against the CPU oracle (oracle/mlp_oracle.py, device rounding points).

Tolerances: the device accumulates bf16 products in fp32 in its own order and
the oracle in float64, so an activation that sits next to a bf16 rounding
boundary can round the other way; relative L2 errors stay near 1e-3.  A wrong
operand layout, a transposed weight or a dropped term gives errors of order 1.
"""
import numpy as np
import pytest

import mlp_oracle as mo

pytestmark = pytest.mark.gpu

IN, OUT = 24, 3


@pytest.fixture(scope="module")
def learned():
    from bullet3_amd import capi, learned as L
    if capi.device_count() < 1:
        pytest.skip("no HIP device")
    return L


def rand_params(rng, in_dim, out_dim, scale=1.0):
    parts = []
    for shp in [(in_dim, 128), (128,), (128, 128), (128,), (128, out_dim), (out_dim,)]:
        fan = shp[0] if len(shp) == 2 else 128
        parts.append(rng.normal(0, scale / np.sqrt(fan), size=shp).ravel())
    return np.concatenate(parts).astype(np.float32)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("rows", [1, 31, 32, 1000, 70000])
def test_forward_matches_oracle(learned, rows):
    rng = np.random.default_rng(rows)
    net = learned.WarmStartNet(IN, OUT, max_train_rows=1024, seed=1)
    flat = rand_params(rng, IN, OUT)
    net.set_params(flat)
    assert np.array_equal(net.get_params(), flat)
    x = rng.normal(size=(rows, IN)).astype(np.float32)
    y = net.forward(x)
    yo = mo.forward(flat, x, IN, OUT, exact=False)
    print("rows %d rel %.2e" % (rows, rel(y, yo)), flush=True)
    assert y.shape == (rows, OUT)
    assert rel(y, yo) < 5e-3
    assert np.max(np.abs(y - yo)) < 2e-2 * max(1.0, np.max(np.abs(yo)))
    assert np.array_equal(net.forward(x), y)   # deterministic
    net.close()


def test_forward_single_unit_columns(learned):
    """one-hot inputs through identity-like weights: pins the lane layout of
    the operands exactly (every product is a single exact term)"""
    net = learned.WarmStartNet(IN, OUT, max_train_rows=64, seed=1)
    p = {k: np.zeros(s, np.float32) for k, s in learned.param_shapes(IN, OUT)}
    for i in range(IN):
        p["w1"][i, i] = 1.0 + i / 32.0          # bf16-exact
    for j in range(128):
        p["w2"][j, 127 - j] = 1.0               # reversal
    for o in range(OUT):
        p["w3"][127 - o, o] = 1.0
        p["w3"][127 - (o + 8), o] = 0.5
    net.set_params(learned.pack(p, IN, OUT))
    x = np.zeros((64, IN), np.float32)
    for r in range(64):
        x[r, r % IN] = 2.0
    y = net.forward(x)
    yo = mo.forward(learned.pack(p, IN, OUT), x, IN, OUT, exact=False)
    assert np.allclose(y, yo, rtol=0, atol=1e-6), np.argwhere(~np.isclose(y, yo, atol=1e-6))[:5]
    net.close()


@pytest.mark.parametrize("rows", [32, 200, 4096])
def test_gradients_match_oracle(learned, rows):
    rng = np.random.default_rng(100 + rows)
    net = learned.WarmStartNet(IN, OUT, max_train_rows=4096, seed=1)
    flat = rand_params(rng, IN, OUT)
    net.set_params(flat)
    x = rng.normal(size=(rows, IN)).astype(np.float32)
    t = rng.normal(size=(rows, OUT)).astype(np.float32)
    loss, g = net.loss_gradients(x, t)
    lo, go = mo.loss_gradients(flat, x, t, IN, OUT, exact=False)
    assert abs(loss - lo) <= 1e-3 * lo
    gp, gop = learned.unpack(g, IN, OUT), mo.unpack(go, IN, OUT)
    for k in ("w1", "b1", "w2", "b2", "w3", "b3"):
        e = rel(gp[k], gop[k])
        print("rows %d %s rel %.2e" % (rows, k, e), flush=True)
        assert e < 1e-2, k
    # and against the unrounded network, looser: bf16 is the only difference
    _, ge = mo.loss_gradients(flat, x, t, IN, OUT, exact=True)
    assert rel(g, ge) < 5e-2
    net.close()


def test_training_fits_a_teacher(learned):
    rng = np.random.default_rng(11)
    teacher = rand_params(rng, IN, OUT, scale=1.0)
    x = rng.normal(size=(20000, IN)).astype(np.float32)
    t = mo.forward(teacher, x, IN, OUT, exact=True).astype(np.float32)
    net = learned.WarmStartNet(IN, OUT, max_train_rows=1024, seed=5)
    y0 = net.forward(x)
    l0 = float(np.mean((y0 - t) ** 2))
    hist = net.train(x, t, batch_rows=1024, steps=600, lr=2e-3, seed=9)
    y1 = net.forward(x)
    l1 = float(np.mean((y1 - t) ** 2))
    print("loss %.4f -> %.4f (last batch %.4f)" % (l0, l1, hist[-1]), flush=True)
    assert np.all(np.isfinite(hist))
    assert l1 < 0.1 * l0
    net.close()


def _mix32(x):
    x = np.uint64(x) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    return x


def test_one_adam_step_matches_oracle(learned):
    """the minibatch rows (hash of seed, step, row) and the Adam update"""
    rng = np.random.default_rng(21)
    n, b, seed, lr = 5000, 256, 77, 1e-3
    net = learned.WarmStartNet(IN, OUT, max_train_rows=b, seed=1)
    flat = rand_params(rng, IN, OUT)
    net.set_params(flat)
    x = rng.normal(size=(n, IN)).astype(np.float32)
    t = rng.normal(size=(n, OUT)).astype(np.float32)
    net.train(x, t, batch_rows=b, steps=1, lr=lr, seed=seed)
    after = net.get_params()
    idx = np.array([(int(_mix32(seed ^ int(_mix32((0 * 0x9E3779B9 + j) & 0xFFFFFFFF)))) * n) >> 32 for j in range(b)])
    _, go = mo.loss_gradients(flat, x[idx], t[idx], IN, OUT, exact=False)
    exp, _, _ = mo.adam(flat.astype(np.float64), go, 0.0, 0.0, 1, lr)
    # first Adam step = lr * sign(g): parameters whose gradient is not tiny agree
    big = np.abs(go) > 1e-3 * np.max(np.abs(go))
    agree = np.abs(after - exp) <= 1e-2 * lr
    print("agree %.4f of %d" % (agree[big].mean(), big.sum()), flush=True)
    assert agree[big].mean() > 0.99
    net.close()


def test_input_normalization_matches_oracle(learned):
    rng = np.random.default_rng(31)
    net = learned.WarmStartNet(IN, OUT, max_train_rows=512, seed=1)
    flat = rand_params(rng, IN, OUT)
    net.set_params(flat)
    shift = rng.normal(size=IN).astype(np.float32) * 3
    scale = np.exp(rng.normal(size=IN)).astype(np.float32)
    net.set_input_normalization(shift, scale)
    s2, c2 = net.get_input_normalization()
    assert np.array_equal(s2, shift) and np.array_equal(c2, scale)
    x = (rng.normal(size=(300, IN)) * 4 + 2).astype(np.float32)
    t = rng.normal(size=(300, OUT)).astype(np.float32)
    y = net.forward(x)
    yo = mo.forward(flat, x, IN, OUT, exact=False, shift=shift, scale=scale)
    assert rel(y, yo) < 5e-3
    loss, g = net.loss_gradients(x, t)
    lo, go = mo.loss_gradients(flat, x, t, IN, OUT, exact=False, shift=shift, scale=scale)
    assert abs(loss - lo) <= 1e-3 * lo
    assert rel(g, go) < 1e-2
    net.close()
