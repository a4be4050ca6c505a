"""Pins for the oracle where no reference golden vector exists: closed forms and
invariants that a plausible slip (a dropped term, a flipped sign, a swapped
index) would break.  CPU only; worlds are built on the host.

Narrowphase closed forms (sat.cl:577-798, satClipHullContacts.cl:720-1000):
  * a unit box pressed delta into the ground (body 0 = A): one manifold, normal
    (0,-1,0) (from B towards A), the MPR point at index 0 and three of the box's
    bottom corners, every depth -delta;
  * any positive gap gives no contact;
  * two boxes side by side pressed delta along x: normal along -x;
  * a box tilted by theta about z: deepest depth = -(delta) at the lowest edge.
PGS closed forms (solveContact.cl:227-303, solveFriction.cl:288-408):
  * a box resting on the ground with depth = -drift (bias 0) and falling
    velocity -g*dt: the normal impulses sum to m*g*dt and the velocity ends at 0;
  * two equal boxes meeting head-on: momentum is conserved exactly and the
    relative normal velocity ends at 0; impulses stay >= 0 and friction within
    mu * sum(normal).
Batching invariants (batchingKernelsNew.cl:144-231): every contact batched, no
dynamic body twice in any 64-contact chunk of a batch, cells partition the set.
Broadphase: grid == SAP == brute force as sets (PairBench.cpp:593-656 protocol).
"""
import numpy as np
import pytest

import pyoracle as po
from bullet3_amd import layout, scenes
from helpers import aligned
from host_world import HostWorld

G = 9.8
DT = 1.0 / 60.0


def box_scene(pos, orn=None, masses=(1.0,), ground=True):
    shapes = [scenes.box_points(100.0, 0.5, 100.0), scenes.box_points(0.5, 0.5, 0.5)]
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    n = len(pos)
    orn = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1)) if orn is None else np.asarray(orn, np.float32)
    mass = np.asarray(masses, np.float32)
    if ground:
        pos = np.concatenate([[[0.0, -0.5, 0.0]], pos]).astype(np.float32)
        orn = np.concatenate([[[0, 0, 0, 1]], orn]).astype(np.float32)
        mass = np.concatenate([[0.0], mass]).astype(np.float32)
        shape = np.array([0] + [1] * n, np.int32)
    else:
        shape = np.ones(n, np.int32)
    return scenes.Scene("closed-form", shapes, mass, pos, orn, shape)


def contacts_of(hw, pairs):
    pairs = np.asarray(pairs, np.int32).reshape(-1, 4)
    n, c, p = po.narrowphase(pairs, hw.bodies, hw.t, capacity=4 * len(pairs) + 4)[:3]
    return c[:n].view(layout.CONTACT).reshape(-1), p


def quat_z(theta):
    return np.array([0.0, 0.0, np.sin(theta / 2), np.cos(theta / 2)], np.float32)


@pytest.mark.parametrize("delta", [0.001, 0.01, 0.019])
def test_box_on_ground(delta):
    x0, z0 = 0.1, 0.2
    hw = HostWorld(box_scene([[x0, 0.5 - delta, z0]]))
    c, p = contacts_of(hw, [[0, 1, -1, -1]])
    assert len(c) == 1 and p[0, 2] == 0
    c = c[0]
    assert c["bodyA"] == 0 and c["bodyB"] == 1
    assert np.allclose(c["worldNormalOnB"][:3], [0.0, -1.0, 0.0], atol=1e-6)
    assert int(c["worldNormalOnB"][3]) == 4
    pts = c["worldPosB"]
    assert np.allclose(pts[:, 3], -delta, atol=2e-6)
    # points 1..3: distinct bottom corners of the box (on B)
    corners = {(round(x0 + sx * 0.5, 5), round(z0 + sz * 0.5, 5)) for sx in (-1, 1) for sz in (-1, 1)}
    got = {(round(float(q[0]), 5), round(float(q[2]), 5)) for q in pts[1:]}
    assert len(got) == 3 and got <= corners
    assert np.allclose(pts[1:, 1], -delta, atol=2e-6)
    # point 0 (MPR) lies inside the footprint, between the two surfaces
    assert abs(pts[0, 0] - x0) <= 0.5 + 1e-6 and abs(pts[0, 2] - z0) <= 0.5 + 1e-6
    assert -delta - 1e-6 <= pts[0, 1] <= 1e-6


@pytest.mark.parametrize("gap", [1e-4, 0.01, 0.3])
@pytest.mark.parametrize("theta", [0.0, 0.4])
def test_gap_gives_no_contact(gap, theta):
    lowest = 0.5 * (np.cos(theta) + np.sin(theta))
    hw = HostWorld(box_scene([[0.0, lowest + gap, 0.0]], orn=[quat_z(theta)]))
    c, p = contacts_of(hw, [[0, 1, -1, -1]])
    assert len(c) == 0 and p[0, 2] == -1


@pytest.mark.parametrize("delta", [0.002, 0.015])
def test_boxes_side_by_side(delta):
    hw = HostWorld(box_scene([[0.0, 0.0, 0.0], [1.0 - delta, 0.0, 0.0]], masses=(1.0, 1.0), ground=False))
    c, _ = contacts_of(hw, [[0, 1, -1, -1]])
    assert len(c) == 1
    n = c[0]["worldNormalOnB"][:3]
    assert np.allclose(np.abs(n), [1.0, 0.0, 0.0], atol=1e-6)
    # normal points from B (at +x) towards A
    assert n[0] < 0
    assert np.allclose(c[0]["worldPosB"][:, 3], -delta, atol=2e-6)


@pytest.mark.parametrize("theta", [0.2, 0.5])
def test_tilted_box_lowest_edge(theta):
    delta = 0.01
    lowest = 0.5 * (np.cos(theta) + np.sin(theta))
    hw = HostWorld(box_scene([[0.0, lowest - delta, 0.0]], orn=[quat_z(theta)]))
    c, _ = contacts_of(hw, [[0, 1, -1, -1]])
    assert len(c) == 1
    assert np.allclose(c[0]["worldNormalOnB"][:3], [0.0, -1.0, 0.0], atol=1e-6)
    d = c[0]["worldPosB"][:int(c[0]["worldNormalOnB"][3]), 3]
    assert abs(d.min() + delta) < 2e-5
    # only the lowest edge penetrates: no point deeper than delta
    assert d.min() >= -delta - 2e-5


def solve(hw, contacts, iters=10, friction=0.7):
    prm = po.solver_params(iterations=iters, friction=friction)
    out = po.solve_contacts(hw.bodies, hw.inertias, contacts, prm, static_idx=0)
    return out["constraints"].view(layout.CONSTRAINT).reshape(-1)


def test_resting_box_impulse_equals_weight():
    drift = 0.005
    m = 2.0
    hw = HostWorld(box_scene([[0.0, 0.5 - drift, 0.0]], masses=(m,)))
    hw.bodies["linVel"][1, 1] = -G * DT
    c, _ = contacts_of(hw, [[0, 1, -1, -1]])
    assert len(c) == 1 and np.allclose(c["worldPosB"][0, :, 3], -drift, atol=1e-6)
    con = solve(hw, c, iters=20)
    lam = con["appliedRambdaDt"][0]
    assert np.all(lam >= 0.0)
    assert abs(lam.sum() - m * G * DT) < 1e-3 * m * G * DT
    assert abs(hw.bodies["linVel"][1, 1]) < 1e-3 * G * DT
    # the static ground keeps zero velocity
    assert np.all(hw.bodies["linVel"][0] == 0.0) and np.all(hw.bodies["angVel"][0] == 0.0)


@pytest.mark.parametrize("delta", [0.01, 0.015])
def test_penetration_bias_pushes_out(delta):
    """b = (depth + drift) * coeff / dt (b3ConvertConstraint4.h:62-148) drives the
    normal velocity of a resting, penetrating box to -b: v = (delta - drift) * 0.2 / dt"""
    drift = 0.005
    hw = HostWorld(box_scene([[0.0, 0.5 - delta, 0.0]], masses=(1.0,)))
    c, _ = contacts_of(hw, [[0, 1, -1, -1]])
    con = solve(hw, c, iters=30)
    want = (delta - drift) * 0.2 / DT
    assert abs(hw.bodies["linVel"][1, 1] - want) < 1e-3 * want
    assert np.all(con["appliedRambdaDt"][0] >= 0.0)


def test_separating_contact_does_not_pull():
    drift = 0.005
    hw = HostWorld(box_scene([[0.0, 0.5 - drift, 0.0]], masses=(1.0,)))
    hw.bodies["linVel"][1, 1] = 1.0
    c, _ = contacts_of(hw, [[0, 1, -1, -1]])
    con = solve(hw, c)
    assert np.all(con["appliedRambdaDt"][0] == 0.0)
    assert hw.bodies["linVel"][1, 1] == np.float32(1.0)


@pytest.mark.parametrize("vx,kinetic", [(1.0, True), (0.05, False)])
def test_friction_bound_on_sliding_box(vx, kinetic):
    """friction impulses are clamped to mu * sum(normal impulses) (solveFriction.cl:309-357):
    a sliding box resting under its weight loses mu*g*dt of speed, a slow one stops"""
    drift = 0.005
    mu = 0.7
    hw = HostWorld(box_scene([[0.0, 0.5 - drift, 0.0]], masses=(1.0,)))
    hw.bodies["linVel"][1, 1] = -G * DT
    hw.bodies["linVel"][1, 0] = vx
    c, _ = contacts_of(hw, [[0, 1, -1, -1]])
    con = solve(hw, c, iters=30, friction=mu)
    lam = con["appliedRambdaDt"][0].sum()
    assert abs(lam - G * DT) < 1e-3 * G * DT
    f = con["fAppliedRambdaDt"][0]
    if kinetic:
        assert abs(np.hypot(f[0], f[1]) - mu * lam) < 1e-6 + 1e-4 * mu * lam
        assert abs(hw.bodies["linVel"][1, 0] - (vx - mu * lam)) < 1e-4
    else:
        assert np.all(np.abs(f) <= mu * lam + 1e-7)
        # the manifold centre (the friction rows' anchor) stops sliding, to 0.1%
        # of the initial speed (the two tangent rows and the twist damping couple)
        r = (con["center"][0] - hw.bodies["pos"][1])[:3].astype(np.float64)
        vc = hw.bodies["linVel"][1, :3] + np.cross(hw.bodies["angVel"][1, :3].astype(np.float64), r)
        assert abs(vc[0]) < 1e-3 * vx and abs(vc[2]) < 1e-3 * vx


def test_head_on_conserves_momentum():
    drift = 0.005
    hw = HostWorld(box_scene([[0.0, 0.0, 0.0], [1.0 - drift, 0.0, 0.0]], masses=(1.5, 1.5), ground=False))
    v = 0.7
    hw.bodies["linVel"][0, 0] = v
    hw.bodies["linVel"][1, 0] = -v
    c, _ = contacts_of(hw, [[0, 1, -1, -1]])
    assert len(c) == 1
    p0 = (hw.bodies["linVel"][:, 0] / hw.bodies["invMass"]).sum()
    con = solve(hw, c, iters=20, friction=0.5)
    p1 = (hw.bodies["linVel"][:, 0] / hw.bodies["invMass"]).sum()
    assert abs(p1 - p0) < 1e-5
    assert abs(hw.bodies["linVel"][0, 0] - hw.bodies["linVel"][1, 0]) < 1e-3 * v
    lam = con["appliedRambdaDt"][0]
    assert np.all(lam >= 0.0)
    bound = 0.5 * lam.sum()
    assert np.all(np.abs(con["fAppliedRambdaDt"][0]) <= bound + 1e-7)


def pile_contacts(steps=30):
    sc = scenes.cube_grid(6, 4, 6, spacing=1.0, drop=0.02, seed=5)
    hw = HostWorld(sc, solver=po.solver_params(iterations=10, friction=0.5))
    for _ in range(steps):
        hw.step()
    n = hw.num_contacts
    return hw, hw.contacts[:n].copy()


def test_batching_invariants():
    hw, contacts = pile_contacts()
    assert len(contacts) > 100
    out = po.solve_contacts(hw.bodies, hw.inertias, contacts, po.solver_params(iterations=4), static_idx=0)
    sc = out["sorted_contacts"].view(layout.CONTACT).reshape(-1)
    counts, offsets, nb = out["counts"].astype(int), out["offsets"].astype(int), out["batch_sizes"]
    assert counts.sum() == len(contacts)
    assert np.array_equal(offsets, np.concatenate([[0], np.cumsum(counts)[:-1]]))
    # same multiset of manifolds
    key = lambda c: sorted(zip(c["bodyA"].tolist(), c["bodyB"].tolist()))
    assert key(sc) == key(contacts)
    inv = hw.bodies["invMass"]
    for cell in range(256):
        seg = sc[offsets[cell]:offsets[cell] + counts[cell]]
        if not len(seg):
            continue
        b = seg["batchIdx"]
        assert np.all(np.diff(b) >= 0) and b[0] == 0 and b[-1] == nb[cell] - 1
        for bi in range(nb[cell]):
            rows = seg[b == bi]
            assert len(rows) > 0
            for s in range(0, len(rows), 64):
                chunk = rows[s:s + 64]
                dyn = [abs(int(x)) for x in np.concatenate([chunk["bodyA"], chunk["bodyB"]]) if inv[abs(int(x))] > 0]
                assert len(dyn) == len(set(dyn)), "cell %d batch %d shares a body" % (cell, bi)


@pytest.mark.parametrize("n,seed", [(2000, 1), (5000, 2)])
def test_broadphase_grid_sap_brute_agree(n, seed):
    rng = np.random.default_rng(seed)
    h = np.exp(rng.uniform(np.log(0.05), np.log(1.0), (n, 1))).astype(np.float32)
    side = (n * 0.539) ** (1.0 / 3.0)
    c = rng.uniform(-side / 2, side / 2, (n, 3)).astype(np.float32)
    a = np.zeros(n, layout.AABB)
    a["min"][:, :3] = c - h
    a["max"][:, :3] = c + h
    a["min"][:, 3] = np.arange(n, dtype=np.int32).view(np.float32)
    small = np.arange(n, dtype=np.int32)
    large = np.zeros(0, np.int32)
    cell = float(2.0 * h.max()) + 1e-3
    g, gt = po.grid_pairs(a, small, large, cell=(cell, cell, cell), max_pairs=40 * n)
    s, st, _ = po.sap_pairs(a, small, large, max_pairs=40 * n)
    b = po.brute_pairs(a, small, large)
    norm = lambda p: sorted(map(tuple, np.sort(np.asarray(p)[:, :2], 1).tolist()))
    assert gt == len(g) and st == len(s)
    assert norm(g) == norm(b) == norm(s)
    assert len(b) > n  # the cloud is dense enough to matter


def test_oracle_sort_is_stable():
    rng = np.random.default_rng(3)
    n = 5000
    kv = np.zeros((n, 2), np.uint32)
    kv[:, 0] = rng.integers(0, 50, n)
    kv[:, 1] = np.arange(n)
    out = np.ascontiguousarray(kv.copy())
    po.lib().b3o_sort_keys(n, po.ptr(out))
    order = np.argsort(kv[:, 0], kind="stable")
    assert np.array_equal(out, kv[order])


def unit_sphere_162():
    import os
    import re
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "unit_sphere_162.h")
    vals = [float(x) for x in re.findall(r"\(float\)(-?[0-9.]+)", open(path).read())]
    d = np.array(vals, np.float32).reshape(-1, 3)
    assert len(d) == 162
    return d


def quat_basis64(q):
    x, y, z, w = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def sat_brute(hw, a, b):
    """minimum overlap over the reference's candidate axes (faces of A, faces of
    B, then edge crosses when |EA|*|EB| <= 162 else the 162 fixed directions),
    in float64; returns (min overlap, overlap function)"""
    t = hw.t
    hulls, world = [], []
    for i in (a, b):
        cv = t["convexes"][t["collidables"][hw.bodies["collidableIdx"][i]]["shapeIndex"]]
        R = quat_basis64(hw.bodies["quat"][i])
        v = t["vertices"][cv["vertexOffset"]:cv["vertexOffset"] + cv["numVertices"], :3].astype(np.float64)
        world.append(v @ R.T + hw.bodies["pos"][i, :3].astype(np.float64))
        fn = t["faces"][cv["faceOffset"]:cv["faceOffset"] + cv["numFaces"]]["plane"][:, :3].astype(np.float64) @ R.T
        ed = t["uniqueEdges"][cv["uniqueEdgesOffset"]:cv["uniqueEdgesOffset"] + cv["numUniqueEdges"], :3]
        hulls.append((fn, ed.astype(np.float64) @ R.T))

    def overlap(ax):
        pa, pb = world[0] @ ax, world[1] @ ax
        return min(pa.max() - pb.min(), pb.max() - pa.min())

    axes = list(hulls[0][0]) + list(hulls[1][0])
    ea, eb = hulls[0][1], hulls[1][1]
    if len(ea) * len(eb) <= 162:
        for e0 in ea:
            for e1 in eb:
                c = np.cross(e0, e1)
                if np.linalg.norm(c) > 1e-6:
                    axes.append(c / np.linalg.norm(c))
    else:
        axes.extend(unit_sphere_162().astype(np.float64))
    return min(overlap(ax) for ax in axes), overlap


@pytest.mark.parametrize("kinds", [(1, 1), (1, 2), (2, 3), (3, 3)])
def test_sat_normal_minimises_overlap(kinds):
    """the SAT normal of a penetrating pair attains the minimum overlap over the
    candidate axes (findSeparatingAxis, sat.cl:577-798; unit-sphere variant
    mpr.cl:180-240) and points from B towards A"""
    rng = np.random.default_rng(11 + kinds[0] * 7 + kinds[1])
    shapes = [scenes.box_points(100.0, 0.5, 100.0), scenes.box_points(0.5, 0.5, 0.5),
              scenes.sphere_points(rng, 16), scenes.sphere_points(rng, 32)]
    checked = 0
    for trial in range(60):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        r = rng.uniform(0.55, 0.95)
        pos = np.array([[0.0, 0.0, 0.0], u * r], np.float32)
        orn = scenes.random_quats(rng, 2).astype(np.float32)
        sc = scenes.Scene("sat", shapes, np.array([1.0, 1.0], np.float32), pos, orn, np.array(kinds, np.int32))
        hw = HostWorld(sc)
        dstar, overlap = sat_brute(hw, 0, 1)
        if dstar < 1e-3:
            continue
        pairs = np.array([[0, 1, -1, -1]], np.int32)
        n, c, p, state, sep = po.narrowphase(pairs, hw.bodies, hw.t, capacity=4, want_sep=True)
        assert state[0] == 1
        s = sep[0, :3].astype(np.float64)
        assert abs(np.linalg.norm(s) - 1.0) < 1e-5
        assert abs(overlap(s) - dstar) < 2e-5, (trial, overlap(s), dstar)
        assert np.dot(s, hw.bodies["pos"][0, :3] - hw.bodies["pos"][1, :3]) >= 0.0
        assert n == 1
        checked += 1
    assert checked >= 20


def hand_batches(pairs, n_bodies):
    """batch ids of hand-built manifolds (bodyA, bodyB) that all fall in one
    solver cell (every body at (1, 1, 1)); body 0 is the static ground"""
    bodies = np.zeros(n_bodies, layout.BODY)
    bodies["pos"][:, :3] = 1.0
    bodies["quat"][:, 3] = 1.0
    bodies["invMass"][1:] = 1.0
    inert = np.zeros(n_bodies, layout.INERTIA)
    inert["initInvInertia"][1:, 0, 0] = inert["initInvInertia"][1:, 1, 1] = inert["initInvInertia"][1:, 2, 2] = 1.0
    c = np.zeros(len(pairs), layout.CONTACT)
    for k, (a, b) in enumerate(pairs):
        c["bodyA"][k], c["bodyB"][k] = a, b
        c["childA"][k] = c["childB"][k] = -1
        c["worldPosB"][k, 0, :3] = 1.0
        c["worldNormalOnB"][k] = (0, 1, 0, 1)
    out = po.solve_contacts(aligned(bodies), aligned(inert), aligned(c), po.solver_params(iterations=0), static_idx=0)
    sc = out["sorted_contacts"].view(layout.CONTACT).reshape(-1)
    cell = int(np.argmax(out["counts"]))
    assert out["counts"][cell] == len(pairs)
    return [(int(x["bodyA"]), int(x["bodyB"]), int(x["batchIdx"])) for x in sc], int(out["batch_sizes"][cell])


def test_batching_hashed_bitmap_aliases_bodies_2048_apart():
    """CreateBatchesNew keeps a 2048-bit map of used bodies indexed by body % 2048
    (batchingKernelsNew.cl:116-140): body 2049 collides with body 1 although
    they are different bodies, so its manifold waits for the next batch"""
    got, nb = hand_batches([(0, 1), (0, 2049), (0, 3)], 2050)
    assert got == [(0, 1, 0), (0, 3, 0), (0, 2049, 1)]
    assert nb == 2
    # without the alias the three are independent: one batch
    got, nb = hand_batches([(0, 1), (0, 2050), (0, 3)], 2051)
    assert got == [(0, 1, 0), (0, 3, 0), (0, 2050, 0)]
    assert nb == 1


def test_batching_map_cleared_after_64_accepted():
    """after every 64 accepted manifolds the map is cleared
    (batchingKernelsNew.cl:212-220): manifold (1, 65) reuses body 1 of the
    first accepted manifold yet joins batch 0, as the 65th accepted"""
    pairs = [(0, k) for k in range(1, 65)] + [(1, 65)]
    got, nb = hand_batches(pairs, 66)
    assert got == [(0, k, 0) for k in range(1, 65)] + [(1, 65, 0)]
    assert nb == 1
    # with only 63 before it the map still holds body 1: next batch
    pairs = [(0, k) for k in range(1, 64)] + [(1, 65)]
    got, nb = hand_batches(pairs, 66)
    assert got == [(0, k, 0) for k in range(1, 64)] + [(1, 65, 1)]
    assert nb == 2


def test_batching_swaps_accepted_to_front():
    """an accepted manifold is swapped with the first not yet batched one
    (batchingKernelsNew.cl:196-206): (2, 3) overtakes the rejected (1, 2)"""
    got, nb = hand_batches([(0, 1), (1, 2), (2, 3)], 4)
    assert got == [(0, 1, 0), (2, 3, 0), (1, 2, 1)]
    assert nb == 2
    # a longer chain: rejected entries rotate behind each accepted one
    got, nb = hand_batches([(1, 2), (2, 3), (3, 4), (4, 5)], 6)
    assert got == [(1, 2, 0), (3, 4, 0), (2, 3, 1), (4, 5, 1)]
    assert nb == 2


def test_oracle_threads_give_the_same_bytes():
    """the all-core oracle (per-pair narrowphase in parallel, compacted in pair
    order; the cells of one parity class in parallel) steps bit for bit like
    the one-thread oracle"""
    sc = scenes.convex_pile(6, 4, 6, spacing=1.0, n_hull_shapes=8, drop=0.0, seed=3)
    a = HostWorld(sc, solver=po.solver_params(iterations=10))
    b = HostWorld(sc, solver=po.solver_params(iterations=10))
    try:
        for step in range(25):
            po.set_threads(1)
            a.step()
            po.set_threads(4)
            b.step()
            assert a.bodies.tobytes() == b.bodies.tobytes(), step
            assert a.num_contacts == b.num_contacts
        assert a.num_contacts > 50
    finally:
        po.set_threads(1)
