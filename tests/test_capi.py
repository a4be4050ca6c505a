"""C ABI checks that need no GPU: the library loads, exports every entry point
include/b3hip.h declares, fails loudly without a device, and registers hull
shapes exactly as the reference's b3ConvexUtility does."""
import ctypes
import os
import re

import numpy as np
import pytest

from bullet3_amd import capi, scenes
from helpers import ROOT, ref_lib


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "b3hip.h")).read()
    return sorted(set(re.findall(r"\b(b3h_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    decl = declared_symbols()
    assert len(decl) > 50
    for name in decl:
        assert hasattr(lib, name), name
    assert set(decl) == set(capi.exported_symbols())


def test_no_device_fails_loudly():
    if capi.device_count() > 0:
        pytest.skip("a device is present")
    cfg = capi.default_config(1000)
    out = ctypes.c_void_p()
    rc = capi.lib().b3h_np_create(ctypes.byref(cfg), None, ctypes.byref(out))
    assert rc == -4
    assert b"no HIP device" in capi.lib().b3h_last_error()


def test_config_defaults_match_reference():
    c = capi.default_config()
    assert c.m_maxConvexBodies == 128 * 1024
    assert c.m_maxBroadphasePairs == 16 * 128 * 1024
    assert c.m_maxContactCapacity == c.m_maxBroadphasePairs
    assert c.m_maxConvexVertices == 8192
    p = capi.default_solver_params()
    assert p.m_numIterations == 4 and abs(p.m_frictionCoeff - 0.7) < 1e-7
    assert abs(p.m_solverDt - 1 / 60.0) < 1e-9 and abs(p.m_batchCellSize - 6.0) < 1e-7


def hull_cases():
    rng = np.random.default_rng(11)
    return [("box", scenes.box_points(0.5, 0.5, 0.5)), ("ground", scenes.box_points(100, 0.5, 100)),
            ("slab", scenes.box_points(2.0, 0.1, 0.7)),
            ("h16", scenes.sphere_points(rng, 16)), ("h32", scenes.sphere_points(rng, 32)),
            ("h64", scenes.sphere_points(rng, 64)),
            ("cloud", rng.uniform(-1, 1, size=(40, 3)).astype(np.float32))]


@pytest.mark.parametrize("name,pts", hull_cases())
def test_hull_invariants(name, pts):
    f = capi.polyhedral_features(pts)
    v, planes = f["vertices"], f["planes"]
    # every hull vertex lies on or behind every face plane, each face on its plane
    d = v @ planes[:, :3].T + planes[:, 3]
    assert d.max() < 1e-5
    for fi, idx in enumerate(f["faces"]):
        assert len(idx) >= 3
        # triangles lie on their plane; a face merged from near-coplanar
        # triangles (normals within the 0.999 weld threshold) only nearly
        tol = 0.05 * np.abs(v).max()
        assert np.abs(d[idx, fi]).max() < tol
    # Euler: V - E + F = 2 on the merged polyhedron
    edges = set()
    for idx in f["faces"]:
        for k in range(len(idx)):
            a, b = int(idx[k]), int(idx[(k + 1) % len(idx)])
            edges.add((min(a, b), max(a, b)))
    used = set(int(i) for idx in f["faces"] for i in idx)
    assert len(used) - len(edges) + len(f["faces"]) == 2
    # unit normals, unit edges
    assert np.allclose(np.linalg.norm(planes[:, :3], axis=1), 1, atol=1e-5)
    assert np.allclose(np.linalg.norm(f["edges"], axis=1), 1, atol=1e-5)


def ref_hull(ref, pts, cap=4096):
    pts = np.ascontiguousarray(pts, np.float32)
    counts = np.zeros(4, np.int32)
    v, p, e = (np.zeros((cap, 4), np.float32) for _ in range(3))
    fc, idx = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    assert ref.ref_hull(capi.ptr(pts), len(pts), cap, capi.ptr(counts), capi.ptr(v), capi.ptr(p), capi.ptr(fc),
                        capi.ptr(idx), capi.ptr(e)) == 0
    nv, nf, ni, ne = counts
    faces, off = [], 0
    for f in range(nf):
        faces.append(idx[off:off + fc[f]].copy())
        off += fc[f]
    return dict(vertices=v[:nv, :3], planes=p[:nf], edges=e[:ne, :3], nfaces=nf, faces=faces)


def canon_dirs(e):
    out = []
    for x in e:
        k = np.argmax(np.abs(x))
        out.append(tuple(np.round(x * np.sign(x[k]), 5)))
    return sorted(out)


@pytest.mark.parametrize("name,pts", hull_cases())
def test_hull_matches_reference_convex_utility(name, pts):
    ref = ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    a = capi.polyhedral_features(pts)
    b = ref_hull(ref, pts)
    assert sorted(map(tuple, np.round(a["vertices"], 6))) == sorted(map(tuple, np.round(b["vertices"], 6)))
    assert len(a["faces"]) == b["nfaces"]
    # faces matched by their vertex sets; the enumeration order differs, and a
    # merged face takes the plane of whichever of its triangles is the merge
    # reference (b3ConvexUtility.cpp:161-164), so merged planes agree to the
    # 0.999 weld threshold and triangle planes to float rounding
    def key(vs, idx):
        return tuple(sorted(tuple(np.round(vs[i], 6)) for i in idx))
    ref_faces = {key(b["vertices"], f): b["planes"][k] for k, f in enumerate(b["faces"])}
    matched = 0
    for k, f in enumerate(a["faces"]):
        pk = key(a["vertices"], f)
        if pk not in ref_faces:
            continue  # a different greedy merge of near-coplanar triangles
        matched += 1
        pr, pm = ref_faces[pk], a["planes"][k]
        if len(f) == 3:
            assert np.abs(pr - pm).max() < 2e-5
        else:
            assert float(np.dot(pr[:3], pm[:3])) > 0.999
    assert matched >= 0.9 * len(a["faces"])
    if matched == len(a["faces"]):
        assert canon_dirs(a["edges"]) == canon_dirs(b["edges"])


def test_public_headers_compile_standalone(tmp_path):
    """include/b3hip.h is valid C and C++ on its own; include/b3hip.hpp is valid C++"""
    import subprocess
    inc = os.path.join(ROOT, "include")
    c = tmp_path / "use.c"
    c.write_text('#include "b3hip.h"\nint main(void){ b3hConfig c; return b3h_config_default(&c); }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", inc, str(c)], check=True)
    cc = tmp_path / "use.cpp"
    cc.write_text('#include "b3hip.hpp"\nint main(){ b3hConfig c; b3h_config_default(&c); '
                  'b3GpuNarrowPhase np(c); return np.isValid() ? 0 : 1; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", inc, str(cc)], check=True)
