"""Parity on the reference's own shape tables and recorded narrowphase launches.

* Shapes registered from precomputed polyhedra -- the reference's
  b3ConvexUtility output stored in tests/golden/ref_hulls.npz -- through
  registerConvexHullShape(b3ConvexUtility*) land in the tables unchanged, equal
  what registering the point clouds produces, and a pile built on them steps
  bit for bit like the oracle.
* The recorded mprPenetrationKernel launches of the reference's kernel test
  (test/OpenCL/AllBullet3Kernels/testExecuteBullet3NarrowphaseKernels.cpp:
  380-424, extracted to tests/golden/mprPenetrationKernel*.npz): loaded as raw
  tables, the device MPR marks exactly 0, 1, 46 and 98 pairs as penetrating, as
  the reference test expects, with the oracle's direction, point and depth bit
  for bit.  The recording holds no face tables; each hull gets the faces of its
  own vertex list's hull (mapped onto that list), and the full narrowphase
  (SAT, clipping, reduction) then equals the oracle's bit for bit on those
  recorded body states.
"""
import ctypes

import numpy as np
import pytest

import pyoracle as po
from bullet3_amd import capi, layout, scenes
from bullet3_amd.world import World
from helpers import OracleWorld, aligned, bits_equal, first_diff, ref_hull_fixture
from test_oracle_golden import CASES, load

pytestmark = pytest.mark.gpu


def pile_polyhedra(scene, fx):
    names = ["ground_pile", "box"] + ["pile_hull%02d" % k for k in range(len(scene.shapes) - 2)]
    polys = [fx[n] for n in names]
    for pts, poly in zip(scene.shapes, polys):
        assert np.array_equal(pts, poly["points"])
    return polys


def test_precomputed_polyhedra_keep_reference_tables():
    fx = ref_hull_fixture()
    sc = scenes.convex_pile(4, 2, 4)
    polys = pile_polyhedra(sc, fx)
    a = World(sc, polyhedra=polys)
    b = World(sc)
    ta, tb = a.shape_tables(), b.shape_tables()
    for k in ("collidables", "convexes", "vertices", "uniqueEdges", "faces", "indices", "localAabbs"):
        assert bits_equal(ta[k], tb[k]), k
    # the hull records point at the reference's tables, in the reference's order
    cv = ta["convexes"]
    for s, poly in enumerate(polys):
        r = cv[s]
        v = ta["vertices"][r["vertexOffset"]:r["vertexOffset"] + r["numVertices"]]
        assert v.view(np.float32).reshape(-1, 4)[:, :3].tobytes() == poly["vertices"].tobytes()
        f = ta["faces"][r["faceOffset"]:r["faceOffset"] + r["numFaces"]]
        assert f["plane"].tobytes() == poly["planes"].tobytes()
        for fi, face in enumerate(f):
            idx = ta["indices"][face["indexOffset"]:face["indexOffset"] + face["numIndices"]]
            assert np.array_equal(idx, poly["faces"][fi])
        e = ta["uniqueEdges"][r["uniqueEdgesOffset"]:r["uniqueEdgesOffset"] + r["numUniqueEdges"]]
        assert e.view(np.float32).reshape(-1, 4)[:, :3].tobytes() == poly["edges"].tobytes()
        # local centre = vertex mean, as the reference computes it
        assert abs(float(r["localCenter"][0]) - float(np.float32(poly["vertices"][:, 0].sum(dtype=np.float32)
                                                                   * np.float32(1.0 / len(v))))) < 1e-6
    a.close()
    b.close()


def test_pipeline_register_convex_polyhedron_matches_narrowphase_entry():
    fx = ref_hull_fixture()
    sc = scenes.cube_grid(2, 2, 2)
    w = World(sc)
    n0 = len(w.shape_tables()["collidables"])
    ci = capi.register_convex_polyhedron(None, dict(fx["pile_hull05"]), pipeline=w.pl)
    assert ci == n0
    t = w.shape_tables()
    assert t["convexes"][ci]["numVertices"] == len(fx["pile_hull05"]["vertices"])
    w.close()


def test_reference_polyhedra_pile_bit_exact():
    fx = ref_hull_fixture()
    sc = scenes.convex_pile(6, 4, 6)
    w = World(sc, polyhedra=pile_polyhedra(sc, fx))
    sp = w.solver_params()
    sp.m_numIterations = 10
    w.set_solver(sp)
    ow = OracleWorld(w)
    for step in range(40):
        w.step()
        ow.step()
        gb = w.bodies()
        assert bits_equal(gb, ow.bodies), "step %d\n%s" % (step, first_diff(gb, ow.bodies))
    assert ow.w.numContacts > 100
    w.close()


def mapped_faces(verts4):
    """faces of the hull of one recorded vertex list, indices mapped onto that
    list (nearest recorded vertex of each hull vertex)"""
    pts = np.ascontiguousarray(verts4[:, :3], np.float32)
    f = capi.polyhedral_features(pts)
    d = np.linalg.norm(f["vertices"][:, None, :].astype(np.float64) - pts[None, :, :], axis=2)
    near = d.argmin(axis=1)
    scale = max(1e-6, float(np.abs(pts).max()))
    assert d.min(axis=1).max() < 1e-3 * scale
    return [near[idx].astype(np.int32) for idx in f["faces"]], f["planes"], f["edges"]


def recorded_tables(d):
    """the recording's tables plus faces and unique edges rebuilt per hull"""
    cvx = d["convexes"].copy()
    ci = cvx.view(np.int32).reshape(-1, 24)
    verts = d["vertices"]
    faces, indices, edges = [], [], []
    for s in range(len(ci)):
        nv, vo = int(ci[s, 19]), int(ci[s, 20])
        fs, planes, ue = mapped_faces(verts[vo:vo + nv])
        ci[s, 17], ci[s, 18] = len(faces), len(fs)
        ci[s, 21], ci[s, 22] = len(edges), len(ue)
        for f, p in zip(fs, planes):
            rec = np.zeros(1, layout.FACE)
            rec["plane"] = p
            rec["indexOffset"], rec["numIndices"] = len(indices), len(f)
            faces.append(rec)
            indices.extend(int(i) for i in f)
        for e in ue:
            edges.append(np.array([e[0], e[1], e[2], 0.0], np.float32))
    return dict(collidables=aligned(d["collidables"]), convexes=aligned(cvx), vertices=aligned(verts),
                uniqueEdges=aligned(np.array(edges, np.float32)), faces=aligned(np.concatenate(faces)),
                indices=aligned(np.array(indices, np.int32)))


class RecordedNarrowphase:
    """a narrowphase holding a recorded launch: raw tables, raw bodies"""

    def __init__(self, d, t):
        nb = len(d["bodies"])
        self.cfg = capi.default_config(max(nb, 1024))
        self.np_ = ctypes.c_void_p()
        capi.call("b3h_np_create", ctypes.byref(self.cfg), None, ctypes.byref(self.np_))
        counts = np.array([len(t["collidables"]), len(t["convexes"]), len(t["vertices"]), len(t["uniqueEdges"]),
                           len(t["faces"]), len(t["indices"]), len(t["collidables"])], np.int32)
        la = aligned(np.zeros((len(t["collidables"]), 8), np.float32))
        capi.call("b3h_np_load_shape_tables", self.np_, capi.ptr(counts), capi.ptr(t["collidables"]),
                  capi.ptr(t["convexes"]), capi.ptr(t["vertices"]), capi.ptr(t["uniqueEdges"]), capi.ptr(t["faces"]),
                  capi.ptr(t["indices"]), capi.ptr(la))
        bodies = d["bodies"].view(layout.BODY).reshape(-1)
        out = ctypes.c_int()
        z3 = np.zeros(3, np.float32)
        for i in range(nb):
            b = bodies[i]
            inv = float(b["invMass"])
            capi.call("b3h_np_register_rigid_body", self.np_, int(b["collidableIdx"]), 1.0 / inv if inv else 0.0,
                      capi.ptr(np.ascontiguousarray(b["pos"][:3])), capi.ptr(np.ascontiguousarray(b["quat"])),
                      capi.ptr(z3), capi.ptr(z3 + 1), 0, ctypes.byref(out))
        capi.call("b3h_np_write_all_bodies_to_gpu", self.np_)
        dev = ctypes.c_void_p()
        capi.call("b3h_np_get_bodies_gpu", self.np_, ctypes.byref(dev), None)
        capi.htod(dev.value, np.ascontiguousarray(d["bodies"]))  # the recorded bytes, unchanged
        self.pairs = capi.DeviceBuffer(16 * len(d["pairs"]))
        capi.htod(self.pairs.ptr, np.ascontiguousarray(d["pairs"], np.int32))
        self.n_pairs = len(d["pairs"])

    def compute(self):
        capi.call("b3h_np_compute_contacts", self.np_, ctypes.c_void_p(self.pairs.ptr), self.n_pairs, None, 0)
        dev = ctypes.c_void_p()
        n = ctypes.c_int()
        capi.call("b3h_np_get_contacts_gpu", self.np_, ctypes.byref(dev), ctypes.byref(n))
        contacts = capi.dtoh(dev.value, 112 * n.value).copy() if n.value else np.zeros(0, np.uint8)
        m = ctypes.c_void_p()
        capi.call("b3h_np_get_mpr_debug", self.np_, ctypes.byref(m))
        mpr = capi.dtoh(m.value, 48 * self.n_pairs, np.float32).reshape(-1, 3, 4).copy()
        pairs = capi.dtoh(self.pairs.ptr, 16 * self.n_pairs, np.int32).reshape(-1, 4).copy()
        return contacts.reshape(-1, 112), mpr, pairs

    def close(self):
        self.pairs.free()
        capi.call("b3h_np_destroy", self.np_)


@pytest.mark.parametrize("name", CASES)
def test_recorded_mpr_launches_on_device(name):
    d = load(name)
    t = recorded_tables(d)
    rn = RecordedNarrowphase(d, t)
    contacts, mpr, pairs = rn.compute()
    hit = mpr[:, 2, 1] == 1.0
    # the reference test's expected appended-contact counts
    assert int(hit.sum()) == int(d["expected_contacts"])
    # the oracle's MPR on the same recorded tables and bodies, bit for bit
    n, oc, op = po.mpr_contacts(d["pairs"], d["bodies"], d["collidables"], d["convexes"], d["vertices"])
    assert n == int(hit.sum())
    assert np.array_equal(np.where(hit)[0], np.where(op[:, 2] >= 0)[0])
    cf = oc.view(np.float32).reshape(-1, 28)
    s = op[hit, 2]
    assert (-mpr[hit, 0, :3]).tobytes() == cf[s, 16:19].tobytes()   # direction = -normal
    assert mpr[hit, 1, :3].tobytes() == cf[s, 0:3].tobytes()        # point
    assert (-mpr[hit, 2, 0]).tobytes() == cf[s, 3].tobytes()        # depth
    # full narrowphase on the recorded bodies, faces rebuilt per hull
    on, ocon, opairs = po.narrowphase(d["pairs"], aligned(d["bodies"]), t)
    assert len(contacts) == on
    assert np.array_equal(pairs[:, 2], opairs[:, 2])
    assert bits_equal(contacts, ocon), first_diff(contacts.view(layout.CONTACT).reshape(-1),
                                                  ocon.view(layout.CONTACT).reshape(-1))
    rn.close()
