"""1000-step statistics.  Stacks diverge chaotically, so long trajectories are
not compared; instead the penetration depths, body speeds and PGS residuals of
1000 steps must stay within stated bounds, on the reference CPU scene (10x10x10
cubes, friction 0.5) and on a 1152-body convex pile (boxes, 16/32-point hulls),
10 PGS iterations, dt 1/60.

Bounds (measured values in brackets; after the settling phase):
  cubes   max penetration < 0.05 m [0.026], 99th pct < 0.03 m [0.021],
          mean speed < 0.25 m/s [0.165, i.e. g*dt added by integration],
          last normal-sweep residual < 0.2 [0.08]
  pile    max penetration < 0.35 m [0.18], 99th pct < 0.15 m [0.08],
          max speed < 15 m/s [4.7], last normal-sweep residual < 0.2 [0.075]
Residual = max |impulse change| / max |impulse| of a sweep."""
import numpy as np
import pytest

from bullet3_amd import scenes
from bullet3_amd.world import World

pytestmark = pytest.mark.gpu


def depths(c):
    n = c["worldNormalOnB"][:, 3].astype(int)
    if not len(c):
        return np.zeros(1)
    return -np.concatenate([c["worldPosB"][i, :n[i], 3] for i in range(len(c))])


CASES = {
    "cubes": (lambda: scenes.cube_grid(10, 10, 10), 0.5, 200,
              dict(maxpen=0.05, p99=0.03, meanv=0.25, maxv=None, res=0.2)),
    "pile": (lambda: scenes.convex_pile(12, 8, 12), None, 300,
             dict(maxpen=0.35, p99=0.15, meanv=None, maxv=15.0, res=0.2)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_thousand_steps_within_bounds(name):
    make, friction, settle, b = CASES[name]
    w = World(make())
    sp = w.solver_params()
    sp.m_numIterations = 10
    if friction is not None:
        sp.m_frictionCoeff = friction
    w.set_solver(sp)
    w.track_residual(True)
    for s in range(1000):
        w.step()
        if (s + 1) % 50 or s + 1 < settle:
            continue
        bodies = w.bodies()
        assert np.isfinite(bodies["pos"]).all() and np.isfinite(bodies["linVel"]).all(), "step %d" % s
        d = depths(w.contacts())
        dyn = bodies["invMass"] > 0
        v = np.linalg.norm(bodies["linVel"][dyn, :3], axis=1)
        r = w.residuals()
        assert d.max() < b["maxpen"], "step %d max penetration %.4f" % (s, d.max())
        assert np.percentile(d, 99) < b["p99"], "step %d p99 penetration %.4f" % (s, np.percentile(d, 99))
        if b["meanv"] is not None:
            assert v.mean() < b["meanv"], "step %d mean speed %.4f" % (s, v.mean())
        if b["maxv"] is not None:
            assert v.max() < b["maxv"], "step %d max speed %.3f" % (s, v.max())
        assert r[sp.m_numIterations - 1] < b["res"], "step %d residual %.3e" % (s, r[sp.m_numIterations - 1])
    w.close()
