"""The oracle against the reference's own known answers and its own C++ code.

* data/unittest_data.zip: recorded mprPenetrationKernel launches whose contact
  counts the reference test asserts (0, 1, 46, 98;
  test/OpenCL/AllBullet3Kernels/testExecuteBullet3NarrowphaseKernels.cpp:396-418),
  extracted to tests/golden/*.npz by tests/golden/make_mpr_fixtures.py.
* oracle/_ref: the reference's dual-compile headers (MPR, integration, world
  AABB, constraint set-up) and its hull builder compiled from /root/reference
  (oracle/ref.mk); those tests are skipped where that build is absent.
"""
import ctypes
import os

import numpy as np
import pytest

import pyoracle as po
from helpers import aligned, ref_lib

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["mprPenetrationKernel60", "mprPenetrationKernel61", "mprPenetrationKernel70", "mprPenetrationKernel128"]


def load(name):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    return {k: aligned(d[k]) for k in d.files}


@pytest.mark.parametrize("name", CASES)
def test_mpr_known_answer_counts(name):
    d = load(name)
    n, c, p = po.mpr_contacts(d["pairs"], d["bodies"], d["collidables"], d["convexes"], d["vertices"])
    assert n == int(d["expected_contacts"])
    assert (p[:, 2] >= 0).sum() == n


@pytest.mark.parametrize("name", CASES[2:])
def test_mpr_matches_reference_cpp(name):
    ref = ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    d = load(name)
    pairs = d["pairs"]
    n = len(pairs)
    res = np.zeros(n, np.int32)
    depth = np.zeros(n, np.float32)
    dirs = np.zeros((n, 4), np.float32)
    pos = np.zeros((n, 4), np.float32)
    cnt = ref.ref_mpr(po.ptr(pairs), n, po.ptr(d["bodies"]), po.ptr(d["collidables"]), po.ptr(d["convexes"]),
                      po.ptr(d["vertices"]), po.ptr(res), po.ptr(depth), po.ptr(dirs), po.ptr(pos))
    nc, c, p = po.mpr_contacts(pairs, d["bodies"], d["collidables"], d["convexes"], d["vertices"])
    assert cnt == nc
    hit = np.where(res == 0)[0]
    assert np.array_equal(hit, np.where(p[:, 2] >= 0)[0])
    cf = c.view(np.float32).reshape(-1, 28)
    s = p[hit, 2]
    # positions and depths: the two builds differ only in rounding order
    assert np.abs(cf[s, 0:3] - pos[hit, :3]).max() < 1e-5
    assert np.abs(-cf[s, 3] - depth[hit]).max() < 2e-4
    # the MPR direction is ill-conditioned at support-vertex near-ties; >= 95% agree to 2e-4
    dn = np.abs(cf[s, 16:19] + dirs[hit, :3]).max(axis=1)
    assert (dn < 2e-4).mean() >= 0.95


def test_integrate_matches_reference_cpp():
    ref = ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(3)
    n = 500
    b = np.zeros((n, 20), np.float32)
    b[:, 0:3] = rng.normal(size=(n, 3))
    q = rng.normal(size=(n, 4))
    b[:, 4:8] = q / np.linalg.norm(q, axis=1, keepdims=True)
    b[:, 8:11] = rng.normal(size=(n, 3))
    b[:, 12:15] = rng.normal(size=(n, 3)) * 3
    b[:10, 12:15] = 300.0
    b[:, 17] = 1.0
    a = aligned(b)
    r = aligned(b)
    po.integrate(a, 1.0 / 60.0, 0.99, (0, -9.8, 0))
    g = np.array([0, -9.8, 0], np.float32)
    ref.ref_integrate(po.ptr(r), n, ctypes.c_float(1.0 / 60.0), ctypes.c_float(0.99), po.ptr(g))
    # the reference C++ build uses libm sinf/cosf; the oracle its own polynomial
    assert np.abs(a - r).max() < 2e-6


def test_world_aabbs_match_reference_cpp():
    ref = ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(4)
    n = 300
    b = np.zeros((n, 20), np.float32)
    b[:, 0:3] = rng.normal(size=(n, 3)) * 10
    q = rng.normal(size=(n, 4))
    b[:, 4:8] = q / np.linalg.norm(q, axis=1, keepdims=True)
    bi = b.view(np.int32)
    bi[:, 16] = np.arange(n) % 3
    b[:, 17] = 1.0
    cols = np.zeros((3, 4), np.int32)
    cols[:, 2] = 3
    cols[:, 3] = np.arange(3)
    la = np.zeros((3, 8), np.float32)
    la[:, 0:3] = -rng.uniform(0.1, 1.0, size=(3, 3))
    la[:, 4:7] = rng.uniform(0.1, 1.0, size=(3, 3))
    b, cols, la = aligned(b), aligned(cols), aligned(la)
    out_r = aligned(np.zeros((n, 8), np.float32))
    ref.ref_world_aabbs(n, po.ptr(b), po.ptr(cols), po.ptr(la), po.ptr(out_r))
    out_o = po.world_aabbs(b, cols, la).view(np.float32).reshape(n, 8)
    assert np.abs(out_o[:, [0, 1, 2, 4, 5, 6]] - out_r[:, [0, 1, 2, 4, 5, 6]]).max() < 1e-5
    assert np.array_equal(out_o[:, 3].view(np.int32), np.arange(n))


def test_constraint_setup_matches_reference_cpp():
    ref = ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(5)
    for trial in range(20):
        bodies = np.zeros((2, 20), np.float32)
        bodies[:, 0:3] = rng.normal(size=(2, 3))
        bodies[:, 7] = 1.0
        bodies[:, 8:11] = rng.normal(size=(2, 3))
        bodies[:, 12:15] = rng.normal(size=(2, 3))
        bodies[:, 17] = [1.0, 0.5] if trial % 2 else [0.0, 2.0]
        ii = np.zeros((2, 2, 3, 4), np.float32)
        for k in range(2):
            d = rng.uniform(0.5, 2.0, 3)
            ii[k, 1, 0, 0], ii[k, 1, 1, 1], ii[k, 1, 2, 2] = d
        c = np.zeros(28, np.float32)
        npts = 1 + trial % 4
        c[0:16] = rng.normal(size=16)
        for p in range(4):
            c[4 * p + 3] = -abs(rng.normal()) * 0.01
        nrm = rng.normal(size=3)
        c[16:19] = nrm / np.linalg.norm(nrm)
        c[19] = npts
        ci = c.view(np.int32)
        ci[21] = 0  # batch index
        ci[22], ci[23] = (0, 1)  # bodyA, bodyB
        ci[24] = ci[25] = -1  # child shapes
        bodies, ii, c = aligned(bodies), aligned(ii), aligned(c)
        out = aligned(np.zeros(44, np.float32))
        ref.ref_set_constraint4(po.ptr(bodies[0]), po.ptr(ii[0, 1]), po.ptr(bodies[1]), po.ptr(ii[1, 1]), po.ptr(c),
                                ctypes.c_float(1 / 60.0), ctypes.c_float(0.005), ctypes.c_float(0.2), po.ptr(out))
        # oracle: full solve with zero iterations returns the constraint
        body_rec = bodies.copy()
        inert = np.zeros((2, 24), np.float32)
        inert[:, 12:24] = ii[:, 1].reshape(2, 12)
        res = po.solve_contacts(aligned(body_rec), aligned(inert), aligned(c.reshape(1, 28)),
                                po.solver_params(iterations=0), static_idx=-1)
        mine = res["constraints"][0].view(np.float32)
        valid = np.ones(44, bool)
        valid[23] = False  # centre.w: the C++ build of the header sums x,y,z only
        valid[28 + npts:32] = False  # m_b of unused rows is left uninitialised by the reference
        valid[40:44] = False  # bodies / batch index / padding (ints)
        if npts == 0:
            valid[24:28] = False
        assert np.abs(mine[valid] - out[valid]).max() < 1e-5 * max(1.0, np.abs(out[valid]).max())


def test_body_registration_matches_reference_cpp():
    """registerPhysicsInstance's broadphase AABB and registerRigidBody's inertia
    record, bit for bit against the reference's own b3Transform / b3Matrix3x3 /
    b3Quaternion code (oracle/_ref), for boxes, hulls and ground shapes at random
    poses and masses"""
    ref = ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    from bullet3_amd import capi
    rng = np.random.default_rng(8)
    f = ctypes.c_float
    for trial in range(2000):
        half = rng.uniform(0.05, 3.0, 3).astype(np.float32)
        off = rng.normal(scale=0.1, size=3).astype(np.float32)
        lmin, lmax = (off - half).astype(np.float32), (off + half).astype(np.float32)
        pos = rng.normal(scale=20, size=3).astype(np.float32)
        q = rng.normal(size=4)
        q = (q / np.linalg.norm(q)).astype(np.float32)
        if trial % 7 == 0:
            q = np.array([0, 0, 0, 1], np.float32)
        mass = 0.0 if trial % 11 == 0 else float(np.exp(rng.uniform(np.log(0.5), np.log(4.0))))
        amin, amax = np.zeros(3, np.float32), np.zeros(3, np.float32)
        inert = aligned(np.zeros(24, np.float32))
        capi.call("b3h_compute_body_registration", po.ptr(lmin), po.ptr(lmax), f(mass), po.ptr(pos), po.ptr(q),
                  po.ptr(amin), po.ptr(amax), po.ptr(inert))
        rmin, rmax = np.zeros(3, np.float32), np.zeros(3, np.float32)
        ref.ref_instance_aabb(po.ptr(lmin), po.ptr(lmax), po.ptr(pos), po.ptr(q), po.ptr(rmin), po.ptr(rmax))
        assert amin.tobytes() == rmin.tobytes() and amax.tobytes() == rmax.tobytes(), trial
        rin = aligned(np.zeros(24, np.float32))
        ref.ref_inertia(f(np.float32(mass)), po.ptr(amin), po.ptr(amax), po.ptr(q), po.ptr(rin))
        # the w lane of each row is padding the reference leaves to its vector code
        cols = [c for c in range(24) if c % 4 != 3]
        assert inert[cols].tobytes() == rin[cols].tobytes(), (trial, inert, rin)
