"""Pins the warm-start oracle (oracle/warm_oracle.py) on the CPU: matching a
step against itself returns its own impulses (friction through the world
vector and back), other body pairs and far points do not match, b3PlaneSpace1
gives an orthonormal frame, and the targets of an unchanged solve are zero."""
import numpy as np

import warm_oracle as wo
from bullet3_amd import layout


def synthetic(n=50, seed=0):
    rng = np.random.default_rng(seed)
    c = np.zeros(n, layout.CONSTRAINT)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    c["linear"][:, :3] = nrm
    c["linear"][:, 3] = 0.7
    for k in range(n):
        npt = 1 + k % 4
        c["worldPos"][k, :npt, :3] = rng.normal(size=(npt, 3)) * 0.5
        c["worldPos"][k, :npt, 3] = -rng.uniform(0, 0.02, npt)
        c["jacCoeffInv"][k, :npt] = -rng.uniform(0.3, 2.0, npt)
        c["appliedRambdaDt"][k, :npt] = rng.uniform(0, 1.0, npt)
        c["center"][k, :3] = c["worldPos"][k, :npt, :3].mean(0)
    c["fJacCoeffInv"] = -rng.uniform(0.3, 2.0, (n, 2))
    c["fAppliedRambdaDt"] = rng.normal(size=(n, 2)) * 0.2
    c["b"] = rng.uniform(0, 1, (n, 4)).astype(np.float32)
    c["bodyA"] = np.arange(n) + 1
    c["bodyB"] = 0
    return c


def test_plane_space_orthonormal():
    rng = np.random.default_rng(1)
    for _ in range(200):
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        p, q = wo.plane_space(n)
        m = np.stack([n, p, q]).astype(np.float64)
        assert np.allclose(m @ m.T, np.eye(3), atol=2e-6)


def test_self_match_returns_own_impulses():
    c = synthetic()
    prev = wo.match(c, c)
    npt = (c["jacCoeffInv"] != 0).sum(1)
    assert np.array_equal(prev[:, 7], npt)
    assert np.all(prev[:, 6] == 1)
    for k in range(len(c)):
        assert np.array_equal(prev[k, :npt[k]], c["appliedRambdaDt"][k, :npt[k]])
    assert np.allclose(prev[:, 4:6], c["fAppliedRambdaDt"], atol=2e-6)
    T = wo.targets(c, prev, 0.1633)
    assert np.max(np.abs(T)) < 1e-4


def test_other_pairs_and_far_points_do_not_match():
    c = synthetic()
    d = c.copy()
    d["bodyB"] = 7                       # no pair in common
    assert np.all(wo.match(c, d)[:, :7] == 0)
    e = c.copy()
    e["worldPos"][:, :, :3] += 0.2       # every point moved beyond the 0.1 tolerance
    prev = wo.match(c, e)
    assert np.all(prev[:, 6] == 1) and np.all(prev[:, :4] == 0)
    f = c.copy()
    f["worldPos"][:, :, :3] += 0.05      # within tolerance: nearest point still itself
    prev = wo.match(c, f)
    npt = (c["jacCoeffInv"] != 0).sum(1)
    for k in range(len(c)):
        assert np.array_equal(prev[k, :npt[k]], c["appliedRambdaDt"][k, :npt[k]])
