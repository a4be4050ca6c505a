"""The C++ classes of include/b3hip.hpp, driven the way the reference's demo
drives b3GpuRigidBodyPipeline (bullet3_amd/csrc/demo_scene.cpp), give the same
bodies, bit for bit, as the same scene stepped through the Python mirror."""
import ctypes
import os

import numpy as np
import pytest

from bullet3_amd import capi, layout, scenes
from bullet3_amd.world import World
from helpers import bits_equal, first_diff

pytestmark = pytest.mark.gpu

DEMO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bullet3_amd", "libb3hip_demo.so")


@pytest.mark.parametrize("kind", ["grid", "sap"])
def test_cpp_demo_matches_python_world(kind):
    capi.lib()
    demo = ctypes.CDLL(DEMO)
    sc = scenes.cube_grid(4, 3, 4, spacing=1.05, drop=0.3, seed=3)
    steps, iters, friction = 30, 10, 0.5
    counts = np.array([len(s) for s in sc.shapes], np.int32)
    verts = np.ascontiguousarray(np.concatenate(sc.shapes), np.float32)
    out = np.zeros(sc.num_bodies, layout.BODY)
    p = capi.ptr
    rc = demo.b3h_demo_run_scene(len(sc.shapes), p(counts), p(verts), sc.num_bodies,
                                 p(np.ascontiguousarray(sc.mass, np.float32)),
                                 p(np.ascontiguousarray(sc.pos, np.float32)),
                                 p(np.ascontiguousarray(sc.orn, np.float32)),
                                 p(np.ascontiguousarray(sc.shape, np.int32)), ctypes.c_float(sc.cell_size),
                                 capi.BROADPHASE_GRID if kind == "grid" else capi.BROADPHASE_SAP, iters,
                                 ctypes.c_float(friction), steps, ctypes.c_float(1.0 / 60.0), p(out))
    assert rc == 0, capi.lib().b3h_last_error()
    w = World(sc, broadphase=kind)
    sp = w.solver_params()
    sp.m_numIterations = iters
    sp.m_frictionCoeff = friction
    w.set_solver(sp)
    for _ in range(steps):
        w.step()
    b = w.bodies()
    assert bits_equal(b, out), first_diff(b, out)
    assert np.any(b["linVel"][1:, 1] != 0.0)
    w.close()
