"""Shared helpers for the parity tests: build the oracle's world description from
a product World (same registered shapes, bodies, proxies and parameters), and
compare record arrays."""
import ctypes
import os

import numpy as np

import pyoracle as po
from bullet3_amd import layout

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libb3ref.so")


def ref_lib():
    """The reference's own C++ (compiled from /root/reference into oracle/_ref), or None."""
    if os.path.exists(REF_LIB):
        return ctypes.CDLL(REF_LIB)
    return None


def ref_hull_fixture():
    """{name: tables} of tests/golden/ref_hulls.npz: the reference's
    b3ConvexUtility output for each point cloud (points, vertices (n,3), planes
    (f,4), faces (list of index arrays), edges (e,3))"""
    d = np.load(os.path.join(ROOT, "tests", "golden", "ref_hulls.npz"))
    out = {}
    for name in d["names"]:
        name = str(name)
        fc, idx = d["face_counts_" + name], d["indices_" + name]
        off = np.concatenate([[0], np.cumsum(fc)])
        out[name] = dict(points=d["pts_" + name], vertices=d["vertices_" + name][:, :3],
                         planes=d["planes_" + name], faces=[idx[off[k]:off[k + 1]] for k in range(len(fc))],
                         edges=d["edges_" + name][:, :3])
    return out


def aligned(a, align=64):
    """copy into a 64-byte aligned buffer (the reference's types are 16-byte aligned)"""
    a = np.ascontiguousarray(a)
    buf = np.empty(a.nbytes + align, np.uint8)
    off = (-buf.ctypes.data) % align
    out = buf[off:off + a.nbytes].view(a.dtype).reshape(a.shape)
    out[...] = a
    return out


def solver_params_oracle(p):
    return po.SolverParams(p.m_solverDt, p.m_numIterations, p.m_frictionCoeff, p.m_positionDrift,
                           p.m_positionConstraintCoeff, p.m_batchCellSize, p.m_angularDamping, 0)


class OracleWorld:
    """The oracle's copy of a product World's state (registered data taken from
    the product, then evolved by the oracle alone)."""

    def __init__(self, world, broadphase=0):
        self.t = world.shape_tables()
        self.bodies = np.ascontiguousarray(world.bodies())
        self.inertias = np.ascontiguousarray(world.inertias())
        self.small, self.large = world.proxy_maps()
        sp = world.solver_params()
        self.solver = solver_params_oracle(sp)
        cfg = world.cfg
        self.max_pairs = cfg.m_maxBroadphasePairs
        self.max_contacts = cfg.m_maxContactCapacity
        n = len(self.bodies)
        self.aabbs = np.zeros(n, layout.AABB)
        self.pairs = np.zeros((self.max_pairs, 4), np.int32)
        self.contacts = np.zeros(self.max_contacts, layout.CONTACT)
        self.resid = np.zeros(2 * max(1, sp.m_numIterations), np.float32)
        w = po.World()
        w.numBodies = n
        w.bodies = self.bodies.ctypes.data
        w.inertias = self.inertias.ctypes.data
        for k, name in (("collidables", "collidables"), ("convexes", "convexes"), ("vertices", "vertices"),
                        ("uniqueEdges", "uniqueEdges"), ("faces", "faces"), ("indices", "indices"),
                        ("localAabbs", "localAabbs")):
            setattr(w, k, self.t[name].ctypes.data)
        w.smallMap = self.small.ctypes.data if len(self.small) else None
        w.nSmall = len(self.small)
        w.largeMap = self.large.ctypes.data if len(self.large) else None
        w.nLarge = len(self.large)
        w.broadphase = broadphase
        cs = world.cell_size
        w.cellSize[0] = w.cellSize[1] = w.cellSize[2] = cs
        for k in range(3):
            w.gridDims[k] = 128
        w.maxBodiesPerCell = 256
        w.maxPairs = self.max_pairs
        w.maxContacts = self.max_contacts
        w.solver = self.solver
        w.gravity[0], w.gravity[1], w.gravity[2] = 0.0, -9.8, 0.0
        w.staticIdx = world.static0_index()
        w.worldAabbs = self.aabbs.ctypes.data
        w.pairs = self.pairs.ctypes.data
        w.contacts = self.contacts.ctypes.data
        w.residuals = self.resid.ctypes.data
        self.joints = np.ascontiguousarray(world.constraints())
        w.constraints = self.joints.ctypes.data if len(self.joints) else None
        w.numConstraints = len(self.joints)
        self.compound = world.compound_tables()
        self.compound_capacity = cfg.m_compoundPairCapacity
        w.compound = po.compound_tables(self.compound, self.compound_capacity)
        self.w = w

    def step(self, dt=1.0 / 60.0):
        po.lib().b3o_step(ctypes.byref(self.w), ctypes.c_float(dt))

    def shapes(self):
        return self.t


def body_mismatch(a, b, fields=("pos", "quat", "linVel", "angVel")):
    """max abs difference per field over xyz(w) of two BODY arrays"""
    return {f: float(np.max(np.abs(a[f] - b[f]))) if len(a) else 0.0 for f in fields}


def bits_equal(a, b):
    """identical bytes (shapes/dtypes of the two record views may differ)"""
    return np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()


def sort_pairs(p):
    p = np.asarray(p)[:, :2].copy()
    lo = np.minimum(p[:, 0], p[:, 1])
    hi = np.maximum(p[:, 0], p[:, 1])
    q = np.stack([lo, hi], 1)
    order = np.lexsort((q[:, 1], q[:, 0]))
    return q[order]


def first_diff(a, b, limit=4):
    """describe where two record arrays differ (first records, differing fields)"""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b).view(a.dtype).reshape(a.shape)
    if a.shape != b.shape:
        return "shape %s vs %s" % (a.shape, b.shape)
    ab = a.view(np.uint8).reshape(len(a), -1)
    bb = b.view(np.uint8).reshape(len(b), -1)
    rows = np.nonzero(np.any(ab != bb, axis=1))[0]
    out = ["%d of %d records differ" % (len(rows), len(a))]
    for r in rows[:limit]:
        if a.dtype.names:
            fs = [f for f in a.dtype.names if a[f][r].tobytes() != b[f][r].tobytes()]
            out.append("rec %d: " % r + "; ".join("%s %s vs %s" % (f, a[f][r], b[f][r]) for f in fs))
        else:
            out.append("rec %d: %s vs %s" % (r, a[r], b[r]))
    return "\n".join(out)
