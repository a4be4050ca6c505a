"""Host-only construction of the oracle's world for CPU tests (no device needed).

Shape tables come from the library's host hull routine (the one behind
registerConvexHullShape, b3GpuNarrowPhase.cpp:298-368); bodies and inertias
follow registerRigidBody (b3GpuNarrowPhase.cpp:816-908) and the registration
AABB of registerPhysicsInstance (b3GpuRigidBodyPipeline.cpp:603-669, margin
0.01), computed by the library's host routine b3h_compute_body_registration.
The local centre is summed in numpy float32 in vertex order, as the library
does.
"""
import ctypes

import numpy as np

import pyoracle as po
from bullet3_amd import capi, layout

SHAPE_CONVEX_HULL = 3


def quat_basis(q):
    x, y, z, w = (np.float64(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def compound_tables_host(kids, laabbs, children, infos, nodes, subtrees):
    """registerCompoundShape's tables (b3GpuNarrowPhase.cpp:370-517) through the
    library's host routine; appends to the table lists, returns (collidable,
    local AABB)"""
    n = len(kids)
    ch = np.zeros(n, layout.CHILD_SHAPE)
    for j, (sk, cpos, corn) in enumerate(kids):
        ch["childPosition"][j, :3] = cpos
        ch["childOrientation"][j] = corn
        ch["shapeIndex"][j] = sk
    child_local = np.array([laabbs[int(sk)] for sk, _, _ in kids], layout.AABB)
    aabb = np.zeros(1, layout.AABB)
    info = np.zeros(1, layout.BVH_INFO)
    nd = np.zeros(2 * n, layout.BVH_NODE)
    st = np.zeros(4 * n + 4, layout.BVH_SUBTREE)
    nn, ns = ctypes.c_int(), ctypes.c_int()
    capi.call("b3h_compute_compound_tables", capi.ptr(ch), n, capi.ptr(child_local), capi.ptr(aabb), capi.ptr(info),
              capi.ptr(nd), len(nd), ctypes.byref(nn), capi.ptr(st), len(st), ctypes.byref(ns))
    info["nodeOffset"] = len(nodes)
    info["subTreeOffset"] = len(subtrees)
    col = np.zeros(1, layout.COLLIDABLE)
    col["numChildShapes"] = n
    col["radius"] = np.array([len(infos)], np.int32).view(np.float32)
    col["shapeType"] = 6
    col["shapeIndex"] = len(children)
    children.extend(ch)
    infos.extend(info)
    nodes.extend(nd[:nn.value])
    subtrees.extend(st[:ns.value])
    return col[0], aabb[0]


def shape_tables(shapes):
    cols, convs, verts, edges, faces, idx, laabbs = [], [], [], [], [], [], []
    children, infos, nodes, subtrees = [], [], [], []
    for s, pts in enumerate(shapes):
        if isinstance(pts, dict) and "compound" in pts:
            col, la = compound_tables_host(pts["compound"], laabbs, children, infos, nodes, subtrees)
            cols.append(col)
            convs.append(np.zeros(1, layout.CONVEX)[0])
            laabbs.append(la)
            continue
        if isinstance(pts, dict):
            # registerSphereShape / registerPlaneShape (b3GpuNarrowPhase.cpp:159-232)
            col = np.zeros(1, layout.COLLIDABLE)[0]
            la = np.zeros(1, layout.AABB)[0]
            if "sphere" in pts:
                r = np.float32(pts["sphere"])
                col["shapeType"], col["shapeIndex"], col["radius"] = 7, 0, r
                la["min"][:3], la["max"][:3] = -r, r
            else:
                fr = np.zeros(1, layout.FACE)[0]
                fr["plane"] = np.asarray(pts["plane"], np.float32)
                col["shapeType"], col["shapeIndex"], col["radius"] = 4, len(faces), np.float32(pts["plane"][3])
                faces.append(fr)
                la["min"][:3], la["max"][:3] = -1e30, 1e30
            cols.append(col)
            convs.append(np.zeros(1, layout.CONVEX)[0])
            laabbs.append(la)
            continue
        f = capi.polyhedral_features(np.asarray(pts, np.float32))
        v = f["vertices"].astype(np.float32)
        cv = np.zeros(1, layout.CONVEX)[0]
        c = np.zeros(3, np.float32)
        for p in v:
            c = c + p
        c = c * np.float32(1.0 / len(v))
        cv["localCenter"][:3] = c
        cv["uniqueEdgesOffset"] = len(edges)
        cv["numUniqueEdges"] = len(f["edges"])
        edges.extend(np.concatenate([f["edges"], np.zeros((len(f["edges"]), 1), np.float32)], 1))
        cv["faceOffset"] = len(faces)
        cv["numFaces"] = len(f["faces"])
        for plane, ind in zip(f["planes"], f["faces"]):
            fr = np.zeros(1, layout.FACE)[0]
            fr["plane"] = plane
            fr["indexOffset"] = len(idx)
            fr["numIndices"] = len(ind)
            idx.extend(int(i) for i in ind)
            faces.append(fr)
        cv["vertexOffset"] = len(verts)
        cv["numVertices"] = len(v)
        verts.extend(np.concatenate([v, np.zeros((len(v), 1), np.float32)], 1))
        col = np.zeros(1, layout.COLLIDABLE)[0]
        col["shapeType"] = SHAPE_CONVEX_HULL
        col["shapeIndex"] = s
        cols.append(col)
        convs.append(cv)
        la = np.zeros(1, layout.AABB)[0]
        la["min"][:3] = v.min(0)
        la["max"][:3] = v.max(0)
        laabbs.append(la)
    return dict(collidables=np.array(cols, layout.COLLIDABLE), convexes=np.array(convs, layout.CONVEX),
                vertices=np.array(verts, np.float32).reshape(-1, 4),
                uniqueEdges=np.array(edges, np.float32).reshape(-1, 4) if edges else np.zeros((1, 4), np.float32),
                faces=np.array(faces, layout.FACE), indices=np.array(idx, np.int32),
                localAabbs=np.array(laabbs, layout.AABB),
                compound=dict(children=np.array(children, layout.CHILD_SHAPE), infos=np.array(infos, layout.BVH_INFO),
                              nodes=np.array(nodes, layout.BVH_NODE), subtrees=np.array(subtrees, layout.BVH_SUBTREE)))


def bodies_and_inertias(scene, tables):
    """body and inertia records as registerPhysicsInstance writes them (the
    library's host routine b3h_compute_body_registration: margin-grown
    transformed local AABB, box inertia of twice its extent)"""
    n = scene.num_bodies
    b = np.zeros(n, layout.BODY)
    inr = np.zeros(n, layout.INERTIA)
    amin, amax = np.zeros(3, np.float32), np.zeros(3, np.float32)
    for i in range(n):
        m = np.float32(scene.mass[i])
        b["pos"][i, :3] = scene.pos[i]
        b["quat"][i] = scene.orn[i]
        b["collidableIdx"][i] = int(scene.shape[i])
        b["invMass"][i] = np.float32(1.0) / m if m else np.float32(0.0)
        b["friction"][i] = 1.0
        la = tables["localAabbs"][int(scene.shape[i])]
        rec = np.zeros(1, layout.INERTIA)
        capi.call("b3h_compute_body_registration", capi.ptr(np.ascontiguousarray(la["min"][:3])),
                  capi.ptr(np.ascontiguousarray(la["max"][:3])), ctypes.c_float(float(m)),
                  capi.ptr(np.ascontiguousarray(scene.pos[i], np.float32)),
                  capi.ptr(np.ascontiguousarray(scene.orn[i], np.float32)), capi.ptr(amin), capi.ptr(amax),
                  capi.ptr(rec))
        inr[i] = rec[0]
    return b, inr


class HostWorld:
    """The oracle's world description built on the host from a Scene."""

    def __init__(self, scene, broadphase=0, solver=None, cell_size=None, max_pairs=None, max_contacts=None):
        self.t = shape_tables(scene.shapes)
        self.bodies, self.inertias = bodies_and_inertias(scene, self.t)
        dyn = np.asarray(scene.mass) > 0
        self.small = np.nonzero(dyn)[0].astype(np.int32)
        self.large = np.nonzero(~dyn)[0].astype(np.int32)
        n = scene.num_bodies
        self.max_pairs = max_pairs or 32 * n + 64
        self.max_contacts = max_contacts or 16 * n + 64
        self.solver = solver or po.solver_params()
        self.aabbs = np.zeros(n, layout.AABB)
        self.pairs = np.zeros((self.max_pairs, 4), np.int32)
        self.contacts = np.zeros(self.max_contacts, layout.CONTACT)
        self.resid = np.zeros(2 * max(1, self.solver.m_numIterations), np.float32)
        w = po.World()
        w.numBodies = n
        w.bodies = self.bodies.ctypes.data
        w.inertias = self.inertias.ctypes.data
        for k in ("collidables", "convexes", "vertices", "uniqueEdges", "faces", "indices", "localAabbs"):
            setattr(w, k, self.t[k].ctypes.data)
        w.smallMap = self.small.ctypes.data if len(self.small) else None
        w.nSmall = len(self.small)
        w.largeMap = self.large.ctypes.data if len(self.large) else None
        w.nLarge = len(self.large)
        w.broadphase = broadphase
        cs = float(cell_size or scene.cell_size)
        w.cellSize[0] = w.cellSize[1] = w.cellSize[2] = cs
        for k in range(3):
            w.gridDims[k] = 128
        w.maxBodiesPerCell = 256
        w.maxPairs = self.max_pairs
        w.maxContacts = self.max_contacts
        w.solver = self.solver
        w.gravity[0], w.gravity[1], w.gravity[2] = 0.0, -9.8, 0.0
        w.staticIdx = int(self.large[0]) if len(self.large) and self.large[0] == 0 else -1
        w.worldAabbs = self.aabbs.ctypes.data
        w.pairs = self.pairs.ctypes.data
        w.contacts = self.contacts.ctypes.data
        w.residuals = self.resid.ctypes.data
        self.compound_capacity = 1024 * 1024
        w.compound = po.compound_tables(self.t["compound"], self.compound_capacity)
        self.w = w

    def step(self, dt=1.0 / 60.0):
        po.lib().b3o_step(ctypes.byref(self.w), ctypes.c_float(dt))

    @property
    def num_contacts(self):
        return self.w.numContacts
