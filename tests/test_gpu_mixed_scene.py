"""Every shape and the joints in one scene, on the GPU against the oracle, bit
for bit: boxes, 16-point hulls, three compound shapes and spheres over a
ground box and a static plane, with point-to-point joints between lattice
neighbours (the mixed-shape benchmark scene at a size the oracle steps in
seconds).  Poses and velocities after every step and the contact count must
equal the oracle's, with the grid and the BVH broadphase."""
import numpy as np
import pytest

from bullet3_amd import scenes
from bullet3_amd.world import World
from helpers import OracleWorld, bits_equal, first_diff

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["grid", "plbvh"])
def test_mixed_shapes_with_joints_bit_exact(kind):
    n = 1000
    sc = scenes.compound_scene(n=n, seed=11, ground_half=(12.0, 0.5, 12.0))
    w = World(sc, broadphase=kind)
    sp = w.solver_params()
    sp.m_numIterations = 10
    w.set_solver(sp)
    joints = scenes.lattice_joints(n, 200)
    for a, b, pa, pb in joints:
        w.add_point2point(a, b, pa, pb)
    ow = OracleWorld(w, broadphase=0 if kind == "grid" else 2)
    assert len(ow.joints) == len(joints)
    child = 0
    for s in range(30):
        w.step()
        ow.step()
        gb = w.bodies()
        assert bits_equal(gb, ow.bodies), "step %d\n%s" % (s, first_diff(gb, ow.bodies))
        c = w.contacts()
        assert len(c) == ow.w.numContacts, "step %d" % s
        child = max(child, int(np.count_nonzero((c["childA"] >= 0) | (c["childB"] >= 0))))
    assert child > 50
    w.close()
