"""C ABI checks that need no GPU: the library loads, exports every entry point
include/b3hip.h declares, fails loudly without a device, and registers hull
shapes exactly as the reference's b3ConvexUtility does."""
import ctypes
import os
import re

import numpy as np
import pytest

from bullet3_amd import capi, scenes
from helpers import ROOT, ref_hull_fixture, ref_lib


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


def tables_equal(a, b):
    """bit-for-bit equality of two shape tables, order included"""
    if len(a["vertices"]) != len(b["vertices"]) or len(a["faces"]) != len(b["faces"]):
        return False
    if a["vertices"][:, :3].tobytes() != b["vertices"][:, :3].tobytes():
        return False
    if any(not np.array_equal(x, y) for x, y in zip(a["faces"], b["faces"])):
        return False
    return a["planes"].tobytes() == b["planes"].tobytes() and a["edges"][:, :3].tobytes() == b["edges"][:, :3].tobytes()


def test_hull_tables_equal_reference_fixture():
    """registerConvexHullShape's tables equal the reference's b3ConvexUtility
    tables byte for byte, in the reference's order (vertices, faces, face
    indices, planes, unique edges), for every shape the benchmark scenes use and
    100 seeded clouds (tests/golden/ref_hulls.npz, written by the reference's
    own code)"""
    fx = ref_hull_fixture()
    assert len(fx) == 201
    for name, r in fx.items():
        a = capi.polyhedral_features(r["points"])
        assert tables_equal(a, r), name


def test_hull_tables_equal_reference_live():
    """the same against the reference's code itself on 600 more clouds,
    degenerate ones (integer grids: coplanar faces, collinear and repeated
    points) included"""
    ref = ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(77)
    for k in range(600):
        n = int(rng.integers(4, 129))
        if k % 3 == 0:
            pts = scenes.sphere_points(rng, n)
        elif k % 3 == 1:
            pts = (rng.normal(size=(n, 3)) * rng.uniform(0.01, 10, 3)).astype(np.float32)
        else:
            pts = (rng.integers(-3, 4, size=(n, 3)) * 0.5).astype(np.float32)
        assert tables_equal(capi.polyhedral_features(pts), ref_hull(ref, pts)), k


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


def test_ctypes_structs_match_the_header(tmp_path):
    """the ctypes mirrors of the ABI structs have the header's size and field
    offsets (a C probe compiled with gcc prints them)"""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    mirrors = {"b3hConfig": capi.Config, "b3hSolverParams": capi.SolverParams,
               "b3hConvexUtility": capi.ConvexUtility}
    lines = []
    for cname, cls in mirrors.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in cls._fields_:
            if f[0].startswith("pad"):
                continue
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    src = tmp_path / "probe.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "b3hip_types.h"\nint main(void){\n%s\nreturn 0;}\n'
                   % "\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run([gcc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.splitlines())
    for cname, cls in mirrors.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f in cls._fields_:
            if f[0].startswith("pad"):
                continue
            assert int(got["%s.%s" % (cname, f[0])]) == getattr(cls, f[0]).offset, (cname, f[0])
