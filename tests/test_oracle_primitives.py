"""The oracle's primitive contacts (primitiveContactsKernel,
src/Bullet3OpenCL/NarrowphaseCollision/kernels/primitiveContacts.cl:793-977),
pinned by closed forms.  Normals point from body B towards body A, as for the
hull pairs (a box on the ground gives (0,-1,0) with the ground as A):

* a sphere (r 0.5) 0.05 into the plane y = 0: one point at the sphere's lowest
  point, depth -0.05, normal (0,-1,0), the plane as body A in either pair order;
* two spheres (r 0.5) 0.9 apart: depth -0.1, normal along A - B, point on B's
  surface; apart by more than 1.0: no contact;
* a box 0.01 into the plane: its four bottom corners, depth -0.01, the plane as
  body A; a box with 8 corners below gets 4 of them (reduction);
* a sphere 0.02 into a box's top face: one point on the face, depth -0.02,
  normal +y (the sphere is A); beside an edge the point is on the edge and the
  normal along the offset.
"""
import numpy as np
import pytest

import pyoracle as po
from bullet3_amd import layout, scenes
from host_world import HostWorld

PLANE = {"plane": (0.0, 1.0, 0.0, 0.0)}
SPHERE = {"sphere": 0.5}
BOX = scenes.box_points(0.5, 0.5, 0.5)


def world(shapes, pos, shape_of, mass, orn=None):
    n = len(pos)
    orn = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1)) if orn is None else np.asarray(orn, np.float32)
    sc = scenes.Scene("prim", shapes, np.asarray(mass, np.float32), np.asarray(pos, np.float32), orn,
                      np.asarray(shape_of, np.int32))
    return HostWorld(sc)


def contacts(hw, pairs):
    pairs = np.asarray(pairs, np.int32).reshape(-1, 4)
    n, c, p = po.narrowphase(pairs, hw.bodies, hw.t, capacity=4 * len(pairs) + 4)[:3]
    return c[:n].view(layout.CONTACT).reshape(-1)


@pytest.mark.parametrize("order", [(0, 1), (1, 0)])
def test_sphere_on_plane(order):
    hw = world([PLANE, SPHERE], [(0, 0, 0), (0.3, 0.45, -0.2)], [0, 1], [0.0, 1.0])
    c = contacts(hw, [order + (-1, -1)])
    assert len(c) == 1
    assert c[0]["bodyA"] == 0 and c[0]["bodyB"] == 1
    assert np.allclose(c[0]["worldNormalOnB"][:3], [0, -1, 0], atol=1e-7)
    assert int(c[0]["worldNormalOnB"][3]) == 1
    assert np.allclose(c[0]["worldPosB"][0], [0.3, -0.05, -0.2, -0.05], atol=1e-6)


@pytest.mark.parametrize("gap,hit", [(0.9, True), (1.01, False)])
def test_sphere_sphere(gap, hit):
    hw = world([SPHERE], [(gap, 2.0, 0.0), (0.0, 2.0, 0.0)], [0, 0], [1.0, 1.0])
    c = contacts(hw, [(0, 1, -1, -1)])
    assert len(c) == (1 if hit else 0)
    if hit:
        assert np.allclose(c[0]["worldNormalOnB"][:3], [1, 0, 0], atol=1e-7)
        assert np.allclose(c[0]["worldPosB"][0], [0.5, 2.0, 0.0, gap - 1.0], atol=1e-6)


@pytest.mark.parametrize("order", [(0, 1), (1, 0)])
def test_box_on_plane(order):
    hw = world([PLANE, BOX], [(0, 0, 0), (1.0, 0.49, 2.0)], [0, 1], [0.0, 1.0])
    c = contacts(hw, [order + (-1, -1)])
    assert len(c) == 1 and c[0]["bodyA"] == 0 and c[0]["bodyB"] == 1
    assert np.allclose(c[0]["worldNormalOnB"][:3], [0, -1, 0], atol=1e-7)
    assert int(c[0]["worldNormalOnB"][3]) == 4
    pts = c[0]["worldPosB"]
    assert np.allclose(pts[:, 1], -0.01, atol=1e-6) and np.allclose(pts[:, 3], -0.01, atol=1e-6)
    corners = {(round(float(x), 4), round(float(z), 4)) for x, z in pts[:, [0, 2]]}
    assert corners == {(0.5, 1.5), (0.5, 2.5), (1.5, 1.5), (1.5, 2.5)}


def test_deep_box_reduced_to_four():
    hw = world([PLANE, BOX], [(0, 0, 0), (0.0, 0.2, 0.0)], [0, 1], [0.0, 1.0])
    c = contacts(hw, [(0, 1, -1, -1)])
    assert int(c[0]["worldNormalOnB"][3]) == 4
    # every corner is below the plane; the deepest (a bottom corner) is kept
    assert np.isclose(c[0]["worldPosB"][:, 3].min(), -0.3, atol=1e-6)


@pytest.mark.parametrize("order", [(0, 1), (1, 0)])
def test_sphere_on_box_face(order):
    hw = world([BOX, SPHERE], [(0, 0, 0), (0.1, 0.98, -0.2)], [0, 1], [0.0, 1.0])
    c = contacts(hw, [order + (-1, -1)])
    assert len(c) == 1 and c[0]["bodyA"] == 1 and c[0]["bodyB"] == 0
    assert np.allclose(c[0]["worldNormalOnB"][:3], [0, 1, 0], atol=1e-6)
    assert np.allclose(c[0]["worldPosB"][0], [0.1, 0.5, -0.2, -0.02], atol=1e-5)


def test_sphere_beside_box_edge():
    # centre 0.3 beyond the top edge x = 0.5 and 0.3 above y = 0.5: distance 0.4243
    hw = world([BOX, SPHERE], [(0, 0, 0), (0.8, 0.8, 0.0)], [0, 1], [0.0, 1.0])
    c = contacts(hw, [(1, 0, -1, -1)])
    assert len(c) == 1
    s = np.sqrt(0.5)
    assert np.allclose(c[0]["worldNormalOnB"][:3], [s, s, 0], atol=1e-5)
    assert np.allclose(c[0]["worldPosB"][0][:3], [0.5, 0.5, 0.0], atol=1e-5)
    assert abs(c[0]["worldPosB"][0][3] - (np.sqrt(0.18) - 0.5)) < 1e-5
