"""The two large workloads at their full size.

* The 131,072-body pile (64 x 32 x 64 lattice of boxes and 16-/32-point hulls
  dropped on a ground box, 10 PGS iterations): after 300 graph-replayed steps
  (the lattice has landed and is piling up), one step equals the oracle's
  step from the same state bit for bit; then 100 more steps keep every body
  finite, the 99th-percentile contact penetration below 40 cm (the 32-layer
  heap presses its lower layers deep with 10 iterations: 24.5 cm measured at
  step 311), no body faster than 30 m/s (free fall from the top of the lattice
  ends near 28 m/s), and the last normal sweep's relative residual finite and
  below 1.
* The 4096 batched scenes (16-512 bodies each on their own grounds): the same
  one-step comparison after 30 steps, then 50 steps of the same properties.
The oracle steps on every host thread (its results do not depend on the
thread count)."""
import os

import numpy as np
import pytest

import pyoracle as po
from bullet3_amd import scenes
from bullet3_amd.world import World
from helpers import OracleWorld, bits_equal, first_diff

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def check_properties(w, label):
    b = w.bodies()
    dyn = b["invMass"] > 0
    for f in ("pos", "quat", "linVel", "angVel"):
        assert np.all(np.isfinite(b[f][dyn])), (label, f)
    speed = np.linalg.norm(b["linVel"][dyn, :3], axis=1)
    assert speed.max() < 30.0, (label, float(speed.max()))
    c = w.contacts()
    depth = -c["worldPosB"][:, :, 3]
    npts = c["worldNormalOnB"][:, 3].astype(int)
    d = np.concatenate([depth[npts > k, k] for k in range(4)]) if len(c) else np.zeros(1)
    p99 = float(np.percentile(d, 99))
    assert p99 < 0.40, (label, p99)
    r = w.residuals()
    k = len(r) // 2
    assert np.isfinite(r[k - 1]) and r[k - 1] < 1.0, (label, r[k - 1])
    return p99, float(speed.max()), float(r[k - 1])


def one_step_against_oracle(w, label):
    po.set_threads(host_threads())
    try:
        ow = OracleWorld(w)
        print(label, "oracle copy made", flush=True)
        w.step(DT)
        ow.step(DT)
        gb = w.bodies()
        assert ow.w.numContacts > 0
        assert bits_equal(gb, ow.bodies), "%s: %s" % (label, first_diff(gb, ow.bodies))
        print(label, "one step bit-exact, contacts", ow.w.numContacts, flush=True)
    finally:
        po.set_threads(1)


def test_pile_131k_at_size():
    sc = scenes.convex_pile(64, 32, 64)
    assert int((sc.mass > 0).sum()) == 131072
    w = World(sc, broadphase="grid")
    sp = w.solver_params()
    sp.m_numIterations = 10
    w.set_solver(sp)
    w.track_residual(True)
    w.step_graph(DT, 300)
    one_step_against_oracle(w, "pile")
    for s in range(10):
        w.step_graph(DT, 9)
        w.step(DT)
        print("pile step", 311 + 10 * s, check_properties(w, "pile step %d" % (311 + 10 * s)), flush=True)
    w.close()


def test_batched_4096_scenes_at_size():
    sc = scenes.batched_scenes(4096)
    w = World(sc, broadphase="grid")
    sp = w.solver_params()
    sp.m_numIterations = 10
    w.set_solver(sp)
    w.set_scenes(sc.meta["scene_offsets"])
    w.track_residual(True)
    w.step_graph(DT, 30)
    one_step_against_oracle(w, "scenes")
    for s in range(5):
        w.step_graph(DT, 9)
        w.step(DT)
        print("scenes step", 41 + 10 * s, check_properties(w, "scenes step %d" % (41 + 10 * s)), flush=True)
    w.close()
