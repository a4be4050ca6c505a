"""Joints on the device (createPoint2PointConstraint / createFixedConstraint,
b3GpuRigidBodyPipeline.h:56-58; solved by b3GpuPgsConstraintSolver::solveJoints
before the contacts, b3GpuRigidBodyPipeline.cpp:361-372) against the oracle, bit
for bit: a chain of boxes hanging from a static beam by point-to-point joints
swinging down onto the ground, welded pairs (fixed joints) in a small pile,
a chain whose joints break (flags equal the oracle's), removal by uid, and
graph-replayed stepping."""
import numpy as np
import pytest

from bullet3_amd import scenes
from bullet3_amd.world import World
from helpers import OracleWorld, bits_equal, first_diff

pytestmark = pytest.mark.gpu


def chain_scene(n_links=8, seed=3):
    shapes = [scenes.box_points(50.0, 0.5, 50.0), scenes.box_points(0.5, 0.5, 0.5),
              scenes.box_points(0.3, 0.1, 0.3)]
    mass = [0.0, 0.0]
    pos = [(0.0, -0.5, 0.0), (0.0, 9.0, 0.0)]
    shape = [0, 2]
    for k in range(n_links):
        mass.append(1.0 + 0.25 * (k % 3))
        pos.append((1.1 * (k + 1), 9.0, 0.0))   # horizontal, 0.1 m apart: the chain swings down
        shape.append(1)
    # a small pile beside it, with welded pairs
    rng = np.random.default_rng(seed)
    for k in range(12):
        mass.append(1.0)
        pos.append((-4.0 + 1.05 * (k % 3), 0.6 + 1.05 * (k // 3), 3.0 + rng.uniform(-0.02, 0.02)))
        shape.append(1)
    n = len(mass)
    orn = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1))
    return scenes.Scene("joints", shapes, np.array(mass, np.float32), np.array(pos, np.float32), orn,
                        np.array(shape, np.int32))


def build(breaking=1e30, fixed_breaking=1e30):
    sc = chain_scene()
    w = World(sc)
    sp = w.solver_params()
    sp.m_numIterations = 10
    w.set_solver(sp)
    # beam (body 1) -> link 0 -> link 1 ...: pivots at the shared corners
    w.add_point2point(1, 2, (0.55, 0.0, 0.0), (-0.55, 0.0, 0.0), breaking)
    for k in range(2, 9):
        w.add_point2point(k, k + 1, (0.55, 0.0, 0.0), (-0.55, 0.0, 0.0), breaking)
    # welded pairs in the pile (side by side: pivot at the shared face centre)
    for a in (10, 13, 16):
        w.add_fixed(a, a + 1, (0.525, 0.0, 0.0), (-0.525, 0.0, 0.0), breaking=fixed_breaking)
    return sc, w


def run_bit_exact(w, steps, check_every=10, graph=False):
    ow = OracleWorld(w)
    assert len(ow.joints) == 11
    for s in range(steps):
        if graph:
            w.step_graph(1.0 / 60.0, 1)
        else:
            w.step()
        ow.step()
        if (s + 1) % check_every == 0:
            gb = w.bodies()
            assert bits_equal(gb, ow.bodies), "step %d: %s" % (s, first_diff(gb, ow.bodies))
    return ow


def test_joints_bit_exact():
    sc, w = build()
    ow = run_bit_exact(w, 90)
    b = w.bodies()
    # the chain hangs: link 0 stays within a link length of the beam's end
    assert np.linalg.norm(b["pos"][2, :3] - (b["pos"][1, :3] + [0.55, 0, 0])) < 0.6
    assert np.all(w.constraints()["flags"] == 1)
    assert ow.w.numContacts > 0
    w.close()


def test_joints_graph_replay_bit_exact():
    sc, w = build()
    run_bit_exact(w, 40, graph=True)
    w.close()


def test_breaking_flags_equal_oracle():
    sc, w = build(breaking=1.0, fixed_breaking=0.02)
    ow = run_bit_exact(w, 40)
    flags = w.constraints()["flags"]
    assert np.array_equal(flags, ow.joints["flags"])
    assert 0 < (flags == 0).sum() < len(flags)
    w.close()


def test_remove_constraint_then_step():
    sc, w = build()
    for _ in range(10):
        w.step()
    uids = w.constraints()["uid"]
    w.remove_constraint(int(uids[3]))
    w.remove_constraint(int(uids[-1]))
    c = w.constraints()
    assert len(c) == 9 and int(uids[3]) not in c["uid"].tolist()
    assert c["uid"][3] == uids[-2]  # the last took the removed one's place
    ow = OracleWorld(w)
    for s in range(30):
        w.step()
        ow.step()
    gb = w.bodies()
    assert bits_equal(gb, ow.bodies), first_diff(gb, ow.bodies)
    w.close()
