"""Steps the 131072-body pile (or, with --scene shapes, the mixed-shape scene)
for profiler runs (rocprofv3 --kernel-trace, --pmc).

    python3 -m bullet3_amd.profile_step --warmup 300 --save-state /tmp/pile300.npy
    rocprofv3 --pmc FETCH_SIZE --kernel-include-regex "k_pgs_level|k_narrowphase|k_np_mpr|copyBuffer" \\
        -d out -o fetch --output-format csv -- \\
        python3 -m bullet3_amd.profile_step --state /tmp/pile300.npy --steps 3 --calibrate-copy-mb 256

The first command reaches the benchmark's state (the pile after its warm-up
steps) without a profiler; the second restores it and takes a few eager steps
(one launch per kernel, no graph), then makes one device-to-device copy of a
known size, whose FETCH_SIZE / WRITE_SIZE readings scale the counters to
bytes.
"""
import argparse
import ctypes
import json

import numpy as np

from . import capi, scenes
from .world import World


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=32)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=0, help="graph-launched steps before --save-state")
    ap.add_argument("--save-state", default=None)
    ap.add_argument("--state", default=None, help="bodies saved by --save-state")
    ap.add_argument("--steps", type=int, default=3, help="eager steps (after --state)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=32, help="narrowphase lanes per pair (64, 32, 16)")
    ap.add_argument("--scene", default="pile", choices=["pile", "shapes"],
                    help="pile: the convex pile of nx x ny x nz bodies; shapes: 32768 bodies of boxes, hulls, "
                         "compounds and spheres on a ground box and a plane, with 2048 point-to-point joints")
    args = ap.parse_args()
    if args.scene == "shapes":
        w = World(scenes.compound_scene(n=32768, seed=11, ground_half=(40.0, 0.5, 40.0)), broadphase="grid")
        for a, b, pa, pb in scenes.lattice_joints(32768, 2048):
            w.add_point2point(a, b, pa, pb)
    else:
        w = World(scenes.convex_pile(args.nx, args.ny, args.nz), broadphase="grid")
    capi.call("b3h_np_set_pairs_per_group", w.np_, args.lanes)
    sp = w.solver_params()
    sp.m_numIterations = args.iters
    w.set_solver(sp)
    if args.warmup:
        w.step_graph(1.0 / 60.0, args.warmup)
        w.sync()
    if args.save_state:
        np.save(args.save_state, w.bodies())
    if args.state:
        w.set_bodies(np.load(args.state))
    for _ in range(args.steps):
        w.step(1.0 / 60.0)
    w.sync()
    out = {"bodies": w.num_bodies, "pairs": len(w.pairs()), "contacts": len(w.contacts())}
    if args.calibrate_copy_mb > 0:
        nbytes = args.calibrate_copy_mb << 20
        a, b = capi.DeviceBuffer(nbytes), capi.DeviceBuffer(nbytes)
        capi.call("b3h_memcpy_dtod", ctypes.c_void_p(b.ptr), ctypes.c_void_p(a.ptr), nbytes)
        capi.call("b3h_synchronize")
        out["calibration_copy_bytes"] = nbytes
        a.free()
        b.free()
    w.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
