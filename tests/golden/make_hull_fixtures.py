"""Reference shape tables for the parity tests of shape registration.

For every point cloud the benchmark scenes register (the ground boxes, the unit
box, the pile's 64 hull shapes, the batched scenes' 32) plus a slab and 100
seeded clouds of 4 to 64 points, this script runs the reference's own
b3ConvexUtility::initializePolyhedralFeatures -- compiled from
/root/reference/src by oracle/ref.mk into oracle/_ref/libb3ref.so -- and stores
its output tables in tests/golden/ref_hulls.npz: vertices (float4), face planes
(float4), index count per face, face indices and unique edges (float4), in the
reference's order.  Data only; no reference code is kept.

Run here (needs oracle/_ref):  python tests/golden/make_hull_fixtures.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def cases():
    from bullet3_amd import scenes
    out = [("ground_cube_grid", scenes.box_points(100.0, 0.5, 100.0)),
           ("ground_pile", scenes.box_points(200.0, 0.5, 200.0)),
           ("ground_batched", scenes.box_points(8.0, 0.5, 8.0)),
           ("box", scenes.box_points(0.5, 0.5, 0.5)),
           ("slab", scenes.box_points(2.0, 0.1, 0.7))]
    # the generators draw their shapes before their bodies, so tiny instances
    # carry the full-size scenes' shape pools
    pile = scenes.convex_pile(1, 1, 3)
    out += [("pile_hull%02d" % k, s) for k, s in enumerate(pile.shapes[2:])]
    batched = scenes.batched_scenes(1)
    out += [("batched_hull%02d" % k, s) for k, s in enumerate(batched.shapes[2:])]
    rng = np.random.default_rng(2024)
    for k in range(100):
        n = int(rng.integers(4, 65))
        if k % 3 == 0:
            pts = scenes.sphere_points(rng, n)
        elif k % 3 == 1:
            pts = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
        else:  # many coplanar, collinear and repeated points
            pts = (rng.integers(-2, 3, size=(n, 3)) * 0.25).astype(np.float32)
        out.append(("cloud%03d" % k, np.ascontiguousarray(pts, np.float32)))
    return out


def ref_tables(ref, pts, cap=4096):
    pts = np.ascontiguousarray(pts, np.float32)
    counts = np.zeros(4, np.int32)
    v, p, e = (np.zeros((cap, 4), np.float32) for _ in range(3))
    fc, idx = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    vp = ctypes.c_void_p
    rc = ref.ref_hull(vp(pts.ctypes.data), len(pts), cap, vp(counts.ctypes.data), vp(v.ctypes.data),
                      vp(p.ctypes.data), vp(fc.ctypes.data), vp(idx.ctypes.data), vp(e.ctypes.data))
    assert rc == 0
    nv, nf, ni, ne = (int(x) for x in counts)
    return v[:nv].copy(), p[:nf].copy(), fc[:nf].copy(), idx[:ni].copy(), e[:ne].copy()


def main():
    ref = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libb3ref.so"))
    data = {}
    names = []
    for name, pts in cases():
        v, p, fc, idx, e = ref_tables(ref, pts)
        names.append(name)
        data["pts_" + name] = pts
        data["vertices_" + name] = v
        data["planes_" + name] = p
        data["face_counts_" + name] = fc
        data["indices_" + name] = idx
        data["edges_" + name] = e
    data["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "ref_hulls.npz"), **data)
    print("wrote %d shapes" % len(names))


if __name__ == "__main__":
    main()
