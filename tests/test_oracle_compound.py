"""Oracle compound narrowphase, pinned against the paths it shares operations
with and against brute force.

- A compound whose one child sits at the body origin has the body's own pose
  (qrot of a zero offset adds zeros; the child orientation is the identity),
  so its contacts against the ground box and against a plane must be those of
  the plain hull at that pose: for the box ground, the SAT normal, point count
  and clipped points (all but point 0, which the hull path replaces with its
  MPR point); for the plane, the whole contact but the pair index.
- Compound x compound child pairs (findCompoundPairsKernel's walk of two
  quantized child trees) must be the pairs of children whose quantized boxes
  overlap in world space (checked in float64 with a margin; only clear cases
  count), with the leaf's child number as the child index.
- Compound x hull lists every child of the compound; both static: nothing.
"""
import numpy as np
import pytest

import pyoracle as po
from bullet3_amd import layout, scenes
from host_world import HostWorld


def one_body_world(shape_kind, pos, orn, ground="box"):
    """ground (box, or a plane at y = 0) plus one body of the given kind:
    'box' (unit box) or 'single' (compound of one unit box at the origin)"""
    shapes = [scenes.box_points(10.0, 0.5, 10.0) if ground == "box" else {"plane": (0.0, 1.0, 0.0, 0.0)},
              scenes.box_points(0.5, 0.5, 0.5), {"compound": [(1, (0.0, 0.0, 0.0), (0, 0, 0, 1))]}]
    sc = scenes.Scene("one", shapes, np.array([0.0, 1.0], np.float32),
                      np.array([[0.0, -0.5 if ground == "box" else 0.0, 0.0], pos], np.float32),
                      np.array([[0, 0, 0, 1], orn], np.float32),
                      np.array([0, 1 if shape_kind == "box" else 2], np.int32), cell_size=3.0)
    return HostWorld(sc)


def contacts_of(hw):
    t = hw.t
    pairs = np.array([[0, 1, 0, 0]], np.int32)
    n, c, _, = po.narrowphase(pairs, hw.bodies, t, capacity=8, compound=t["compound"], compound_capacity=64)
    return c[:n].view(layout.CONTACT).reshape(-1)


@pytest.mark.parametrize("seed", range(12))
def test_single_child_compound_on_box_equals_hull(seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    pos = (rng.uniform(-1, 1), rng.uniform(0.3, 0.49), rng.uniform(-1, 1))
    hb = contacts_of(one_body_world("box", pos, q))
    hc = contacts_of(one_body_world("single", pos, q))
    assert len(hb) == len(hc) == 1
    b, c = hb[0], hc[0]
    assert c["childA"] == -1 and c["childB"] == 0
    n = int(c["worldNormalOnB"][3])
    assert n == int(b["worldNormalOnB"][3]) and n >= 1
    assert c["worldNormalOnB"][:3].tobytes() == b["worldNormalOnB"][:3].tobytes()
    for k in range(1, n):
        assert c["worldPosB"][k].tobytes() == b["worldPosB"][k].tobytes()
    # the hull path's point 0 is the MPR point; the compound path keeps the clip point
    assert c["worldPosB"][0][3] <= 0.02


@pytest.mark.parametrize("seed", range(8))
def test_single_child_compound_on_plane_equals_hull(seed):
    rng = np.random.default_rng(100 + seed)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    pos = (rng.uniform(-1, 1), rng.uniform(0.2, 0.49), rng.uniform(-1, 1))
    hb = contacts_of(one_body_world("box", pos, q, ground="plane"))
    hc = contacts_of(one_body_world("single", pos, q, ground="plane"))
    assert len(hb) == len(hc) == 1
    b, c = hb[0].copy(), hc[0].copy()
    assert c["batchIdx"] == 0
    b["batchIdx"] = c["batchIdx"]
    assert b.tobytes() == c.tobytes()


def quat_rows(q):
    x, y, z, w = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def leaf_boxes(t, info, pos, orn):
    """world boxes (float64) of a compound's leaves, by leaf child number"""
    nodes = t["nodes"][info["nodeOffset"]:info["nodeOffset"] + 2 * info["numNodes"] - 1]
    q, mn0 = info["quantization"][:3].astype(np.float64), info["aabbMin"][:3].astype(np.float64)
    R = np.abs(quat_rows(orn))
    out = {}
    for nd in nodes:
        if nd["escapeOrIndex"] < 0:
            continue
        lo = nd["qmin"].astype(np.float64) / q + mn0
        hi = nd["qmax"].astype(np.float64) / q + mn0
        c = quat_rows(orn) @ (0.5 * (lo + hi)) + np.asarray(pos, np.float64)
        e = R @ (0.5 * (hi - lo))
        out[int(nd["escapeOrIndex"])] = (c - e, c + e)
    return out


@pytest.mark.parametrize("seed", range(10))
def test_compound_compound_child_pairs_brute_force(seed):
    rng = np.random.default_rng(200 + seed)
    shapes = [scenes.box_points(10.0, 0.5, 10.0), scenes.box_points(0.2, 0.2, 0.2)]
    for _ in range(2):
        k = int(rng.integers(3, 40))
        kids = []
        for j in range(k):
            q = rng.normal(size=4)
            q /= np.linalg.norm(q)
            kids.append((1, tuple(rng.uniform(-1.5, 1.5, 3)), tuple(q)))
        shapes.append({"compound": kids})
    qa, qb = (v / np.linalg.norm(v) for v in rng.normal(size=(2, 4)))
    pa = np.zeros(3)
    pb = rng.uniform(-1.8, 1.8, 3)
    sc = scenes.Scene("cc", shapes, np.array([0.0, 1.0, 1.0], np.float32),
                      np.array([[0, -20.0, 0], pa, pb], np.float32), np.array([[0, 0, 0, 1], qa, qb], np.float32),
                      np.array([0, 2, 3], np.int32), cell_size=3.0)
    hw = HostWorld(sc)
    t = hw.t
    pairs = np.array([[1, 2, 0, 0]], np.int32)
    res = po.narrowphase(pairs, hw.bodies, t, capacity=4096, compound=t["compound"], compound_capacity=100000,
                         want_child_pairs=True)
    cps, total = res[3], res[4]
    assert total == len(cps)
    got = set((int(r[2]), int(r[3])) for r in cps)
    assert all(r[0] == 1 and r[1] == 2 for r in cps)
    ia, ib = t["compound"]["infos"][0], t["compound"]["infos"][1]
    ba = leaf_boxes(t["compound"], ia, pa, qa)
    bb = leaf_boxes(t["compound"], ib, pb, qb)
    margin = 1e-4
    clear_in, clear_out = set(), set()
    for i, (alo, ahi) in ba.items():
        for j, (blo, bhi) in bb.items():
            gap = np.max(np.maximum(alo - bhi, blo - ahi))
            if gap < -margin:
                clear_in.add((i, j))
            elif gap > margin:
                clear_out.add((i, j))
    assert clear_in <= got
    assert not (got & clear_out)
    # each body pair's child pairs are sorted by (childA, childB)
    keys = [(int(r[2]) & 0xffffffff, int(r[3]) & 0xffffffff) for r in cps]
    assert keys == sorted(keys)


def test_compound_hull_lists_every_child_and_statics_nothing():
    sc = scenes.compound_scene(n=10, plane=False, sphere=False)
    hw = HostWorld(sc)
    t = hw.t
    cols = t["collidables"]
    b = hw.bodies
    comp = [i for i in range(len(b)) if cols[b["collidableIdx"][i]]["shapeType"] == 6]
    hull = [i for i in range(len(b)) if cols[b["collidableIdx"][i]]["shapeType"] == 3 and b["invMass"][i] > 0]
    a, h = comp[0], hull[0]
    pairs = np.array([[a, h, 0, 0], [h, a, 0, 0], [0, a, 0, 0]], np.int32)
    res = po.narrowphase(pairs, b, t, capacity=64, compound=t["compound"], compound_capacity=1000,
                         want_child_pairs=True)
    cps = res[3]
    col = cols[b["collidableIdx"][a]]
    nch, first = int(col["numChildShapes"]), int(col["shapeIndex"])
    want = [(a, h, first + c, -1) for c in range(nch)] + [(h, a, -1, first + c) for c in range(nch)] + \
           [(0, a, -1, first + c) for c in range(nch)]
    assert [tuple(int(v) for v in r) for r in cps] == want
    # both static: no child pairs
    b2 = b.copy()
    b2["invMass"][a] = 0.0
    res = po.narrowphase(np.array([[0, a, 0, 0]], np.int32), b2, t, capacity=8, compound=t["compound"],
                         compound_capacity=100, want_child_pairs=True)
    assert res[4] == 0
