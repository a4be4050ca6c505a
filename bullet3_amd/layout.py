"""numpy views of the b3hip record layouts (include/b3hip_types.h).

Every dtype has the byte layout of the corresponding Bullet3 shared struct, so a
buffer read from the device can be inspected field by field.
"""
import numpy as np

F4 = (np.float32, (4,))

BODY = np.dtype([("pos", F4), ("quat", F4), ("linVel", F4), ("angVel", F4),
                 ("collidableIdx", np.int32), ("invMass", np.float32),
                 ("restitution", np.float32), ("friction", np.float32)])
INERTIA = np.dtype([("invInertiaWorld", np.float32, (3, 4)), ("initInvInertia", np.float32, (3, 4))])
COLLIDABLE = np.dtype([("numChildShapes", np.int32), ("radius", np.float32), ("shapeType", np.int32),
                       ("shapeIndex", np.int32)])
CONVEX = np.dtype([("localCenter", F4), ("extents", F4), ("mC", F4), ("mE", F4), ("radius", np.float32),
                   ("faceOffset", np.int32), ("numFaces", np.int32), ("numVertices", np.int32),
                   ("vertexOffset", np.int32), ("uniqueEdgesOffset", np.int32), ("numUniqueEdges", np.int32),
                   ("unused", np.int32)])
FACE = np.dtype([("plane", F4), ("indexOffset", np.int32), ("numIndices", np.int32), ("pad", np.int32, (2,))])
AABB = np.dtype([("min", F4), ("max", F4)])
CONTACT = np.dtype([("worldPosB", np.float32, (4, 4)), ("worldNormalOnB", F4), ("restitutionCmp", np.uint16),
                    ("frictionCmp", np.uint16), ("batchIdx", np.int32), ("bodyA", np.int32), ("bodyB", np.int32),
                    ("childA", np.int32), ("childB", np.int32), ("unused", np.int32, (2,))])
CONSTRAINT = np.dtype([("linear", F4), ("worldPos", np.float32, (4, 4)), ("center", F4),
                       ("jacCoeffInv", np.float32, (4,)), ("b", np.float32, (4,)),
                       ("appliedRambdaDt", np.float32, (4,)), ("fJacCoeffInv", np.float32, (2,)),
                       ("fAppliedRambdaDt", np.float32, (2,)), ("bodyA", np.uint32), ("bodyB", np.uint32),
                       ("batchIdx", np.int32), ("paddings", np.uint32)])
PAIR = np.dtype([("x", np.int32), ("y", np.int32), ("z", np.int32), ("w", np.int32)])
CONSTRAINT_JOINT = np.dtype([("constraintType", np.int32), ("rbA", np.int32), ("rbB", np.int32),
                             ("breakingImpulseThreshold", np.float32), ("pivotInA", F4), ("pivotInB", F4),
                             ("relTargetAB", F4), ("flags", np.int32), ("uid", np.int32), ("padding", np.int32, (2,))])
RAYINFO = np.dtype([("from", F4), ("to", F4)])
RAYHIT = np.dtype([("hitFraction", np.float32), ("hitBody", np.int32), ("hitResult1", np.int32),
                   ("hitResult2", np.int32), ("hitPoint", F4), ("hitNormal", F4)])
CHILD_SHAPE = np.dtype([("childPosition", F4), ("childOrientation", F4), ("shapeIndex", np.int32),
                        ("numChildShapes", np.int32), ("collidableShapeIndex", np.int32), ("shapeType", np.int32)])
BVH_NODE = np.dtype([("qmin", np.uint16, (3,)), ("qmax", np.uint16, (3,)), ("escapeOrIndex", np.int32)])
BVH_SUBTREE = np.dtype([("qmin", np.uint16, (3,)), ("qmax", np.uint16, (3,)), ("rootNodeIndex", np.int32),
                        ("subtreeSize", np.int32), ("padding", np.int32, (3,))])
BVH_INFO = np.dtype([("aabbMin", F4), ("aabbMax", F4), ("quantization", F4), ("numNodes", np.int32),
                     ("numSubTrees", np.int32), ("nodeOffset", np.int32), ("subTreeOffset", np.int32)])

assert BODY.itemsize == 80 and INERTIA.itemsize == 96 and COLLIDABLE.itemsize == 16
assert CONVEX.itemsize == 96 and FACE.itemsize == 32 and AABB.itemsize == 32
assert CONTACT.itemsize == 112 and CONSTRAINT.itemsize == 176 and PAIR.itemsize == 16
assert RAYINFO.itemsize == 32 and RAYHIT.itemsize == 48 and CONSTRAINT_JOINT.itemsize == 80
assert CHILD_SHAPE.itemsize == 48 and BVH_NODE.itemsize == 16 and BVH_SUBTREE.itemsize == 32 and BVH_INFO.itemsize == 64


def aabb_handle(aabbs):
    """body index stored in min.w (m_minIndices[3])"""
    return aabbs["min"][:, 3].view(np.int32) if aabbs.dtype == AABB else None
