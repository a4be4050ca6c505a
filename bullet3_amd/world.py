"""Python mirror of the reference's demo set-up (GpuRigidBodyDemo::initPhysics,
examples/OpenCL/rigidbody/GpuRigidBodyDemo.cpp:110-152, and the scene
registration of GpuConvexScene.cpp:171-304) over the C ABI:

    np = b3GpuNarrowPhase(config); bp = b3Gpu{Grid,Sap}Broadphase();
    pipeline = b3GpuRigidBodyPipeline(np, bp, config)
    registerConvexHullShape per shape, registerPhysicsInstance per body,
    writeAllInstancesToGpu, then stepSimulation(dt) per frame.

Every call goes through bullet3_amd/libb3hip.so.
"""
import ctypes

import numpy as np

from . import capi
from .capi import call, ptr
from . import layout


class World:
    def __init__(self, scene, broadphase="grid", max_bodies=None, solver=None, grid_dims=(128, 128, 128),
                 max_bodies_per_cell=256, cell_size=None, pairs_per_body=16, polyhedra=None):
        """polyhedra: optional list (one per scene shape) of precomputed tables
        {vertices (n,3), planes (f,4), faces [index arrays], edges (e,3)},
        registered as given (registerConvexHullShape(b3ConvexUtility*))
        instead of hulling the scene's point clouds."""
        n = scene.num_bodies
        if max_bodies is None:
            max_bodies = max(128 * 1024, n + 10)
        self.cfg = capi.default_config(max_bodies, pairs_per_body)
        self.scene = scene
        self.np_ = ctypes.c_void_p()
        self.bp = ctypes.c_void_p()
        self.pl = ctypes.c_void_p()
        call("b3h_np_create", ctypes.byref(self.cfg), None, ctypes.byref(self.np_))
        kind = capi.BROADPHASE_KINDS[broadphase]
        call("b3h_bp_create", kind, ctypes.byref(self.cfg), None, ctypes.byref(self.bp))
        cs = float(cell_size if cell_size is not None else scene.cell_size)
        self.cell_size = cs
        cell3 = np.array([cs, cs, cs], np.float32)
        dims3 = np.array(grid_dims, np.int32)
        call("b3h_bp_set_grid_params", self.bp, ptr(cell3), ptr(dims3), int(max_bodies_per_cell))
        call("b3h_pipeline_create", self.np_, self.bp, ctypes.byref(self.cfg), None, ctypes.byref(self.pl))
        self.collidables = []
        for k, pts in enumerate(scene.shapes):
            if isinstance(pts, dict):
                # primitive shapes: {"sphere": radius} or {"plane": (nx, ny, nz, constant)};
                # compounds: {"compound": [(shape k of an earlier hull, (x, y, z), (qx, qy, qz, qw)), ...]}
                out = ctypes.c_int(-1)
                if "compound" in pts:
                    kids = pts["compound"]
                    ch = np.zeros(len(kids), layout.CHILD_SHAPE)
                    for j, (sk, cpos, corn) in enumerate(kids):
                        ch["childPosition"][j, :3] = cpos
                        ch["childOrientation"][j] = corn
                        ch["shapeIndex"][j] = self.collidables[sk]
                    call("b3h_np_register_compound_shape", self.np_, ptr(ch), len(ch), ctypes.byref(out))
                elif "sphere" in pts:
                    call("b3h_np_register_sphere_shape", self.np_, ctypes.c_float(pts["sphere"]), ctypes.byref(out))
                else:
                    normal = np.ascontiguousarray(pts["plane"][:3], np.float32)
                    call("b3h_np_register_plane_shape", self.np_, ptr(normal), ctypes.c_float(pts["plane"][3]),
                         ctypes.byref(out))
                self.collidables.append(out.value)
                continue
            if polyhedra is not None:
                self.collidables.append(capi.register_convex_polyhedron(self.np_, polyhedra[k]))
                continue
            pts = np.ascontiguousarray(pts, np.float32)
            out = ctypes.c_int(-1)
            call("b3h_np_register_convex_hull_shape", self.np_, ptr(pts), 12, len(pts),
                 ptr(np.ones(3, np.float32)), ctypes.byref(out))
            self.collidables.append(out.value)
        pos = np.ascontiguousarray(scene.pos, np.float32)
        orn = np.ascontiguousarray(scene.orn, np.float32)
        out = ctypes.c_int(-1)
        reg = capi.lib().b3h_pipeline_register_physics_instance
        for i in range(n):
            rc = reg(self.pl, ctypes.c_float(float(scene.mass[i])), ptr(pos[i]), ptr(orn[i]),
                     self.collidables[int(scene.shape[i])], i, 0, ctypes.byref(out))
            capi.check(rc, "registerPhysicsInstance")
        call("b3h_pipeline_write_all_instances_to_gpu", self.pl)
        if solver is not None:
            self.set_solver(solver)

    # -- parameters
    def solver_params(self):
        p = capi.SolverParams()
        call("b3h_pipeline_get_solver_params", self.pl, ctypes.byref(p))
        return p

    def set_solver(self, p):
        call("b3h_pipeline_set_solver_params", self.pl, ctypes.byref(p))

    # -- stepping
    def step(self, dt=1.0 / 60.0):
        call("b3h_pipeline_step_simulation", self.pl, ctypes.c_float(dt))

    def step_graph(self, dt=1.0 / 60.0, count=1):
        call("b3h_pipeline_step_graph", self.pl, ctypes.c_float(dt), int(count))

    def sync(self):
        call("b3h_synchronize")

    # -- state access
    @property
    def num_bodies(self):
        n = ctypes.c_int()
        call("b3h_pipeline_get_num_bodies", self.pl, ctypes.byref(n))
        return n.value

    def bodies(self):
        out = np.zeros(self.num_bodies, layout.BODY)
        call("b3h_np_readback_all_bodies_to_cpu", self.np_, ptr(out))
        return out

    def set_bodies(self, bodies):
        d = ctypes.c_void_p()
        call("b3h_np_get_bodies_gpu", self.np_, ctypes.byref(d), None)
        capi.htod(d.value, np.ascontiguousarray(bodies))

    def compound_tables(self):
        """host copies of the compound tables: children, infos, nodes, subtrees"""
        counts = np.zeros(4, np.int32)
        call("b3h_np_get_compound_tables", self.np_, ptr(counts), None, None, None, None)
        t = dict(children=np.zeros(max(1, counts[0]), layout.CHILD_SHAPE),
                 infos=np.zeros(max(1, counts[1]), layout.BVH_INFO),
                 nodes=np.zeros(max(1, counts[2]), layout.BVH_NODE),
                 subtrees=np.zeros(max(1, counts[3]), layout.BVH_SUBTREE))
        call("b3h_np_get_compound_tables", self.np_, ptr(counts), ptr(t["children"]), ptr(t["infos"]),
             ptr(t["nodes"]), ptr(t["subtrees"]))
        if counts[0] == 0:
            t["children"] = t["children"][:0]
        return t

    def inertias(self):
        out = np.zeros(self.num_bodies, layout.INERTIA)
        call("b3h_np_get_inertias_cpu", self.np_, ptr(out), len(out))
        return out

    def shape_tables(self):
        counts = np.zeros(7, np.int32)
        call("b3h_np_get_shape_tables", self.np_, ptr(counts), None, None, None, None, None, None, None)
        t = dict(collidables=np.zeros(counts[0], layout.COLLIDABLE), convexes=np.zeros(counts[1], layout.CONVEX),
                 vertices=np.zeros((counts[2], 4), np.float32), uniqueEdges=np.zeros((max(counts[3], 1), 4), np.float32),
                 faces=np.zeros(max(counts[4], 1), layout.FACE), indices=np.zeros(max(counts[5], 1), np.int32),
                 localAabbs=np.zeros(counts[6], layout.AABB))
        call("b3h_np_get_shape_tables", self.np_, ptr(counts), ptr(t["collidables"]), ptr(t["convexes"]),
             ptr(t["vertices"]), ptr(t["uniqueEdges"]), ptr(t["faces"]), ptr(t["indices"]), ptr(t["localAabbs"]))
        return t

    def proxy_maps(self):
        c = np.zeros(2, np.int32)
        call("b3h_bp_get_proxy_maps", self.bp, ptr(c), None, None)
        sm = np.zeros(max(c[0], 1), np.int32)
        lm = np.zeros(max(c[1], 1), np.int32)
        call("b3h_bp_get_proxy_maps", self.bp, ptr(c), ptr(sm), ptr(lm))
        return sm[:c[0]].copy(), lm[:c[1]].copy()

    def world_aabbs(self):
        d = ctypes.c_void_p()
        n = ctypes.c_int()
        call("b3h_bp_get_aabb_buffer_ws", self.bp, ctypes.byref(d), ctypes.byref(n))
        return capi.dtoh(d.value, n.value * 32, layout.AABB)

    def pairs(self):
        n = ctypes.c_int()
        call("b3h_bp_get_num_overlap", self.bp, ctypes.byref(n))
        d = ctypes.c_void_p()
        call("b3h_bp_get_overlapping_pair_buffer", self.bp, ctypes.byref(d))
        return capi.dtoh(d.value, n.value * 16, np.int32).reshape(-1, 4)

    def contacts(self):
        d = ctypes.c_void_p()
        n = ctypes.c_int()
        call("b3h_np_get_contacts_gpu", self.np_, ctypes.byref(d), ctypes.byref(n))
        return capi.dtoh(d.value, n.value * 112, layout.CONTACT)

    def solver_debug(self):
        sc, cs, cc, co, bs = (ctypes.c_void_p() for _ in range(5))
        n = ctypes.c_int()
        call("b3h_pipeline_get_solver_debug", self.pl, ctypes.byref(sc), ctypes.byref(cs), ctypes.byref(cc),
             ctypes.byref(co), ctypes.byref(bs), ctypes.byref(n))
        k = n.value
        return dict(sorted_contacts=capi.dtoh(sc.value, k * 112, layout.CONTACT),
                    constraints=capi.dtoh(cs.value, k * 176, layout.CONSTRAINT),
                    counts=capi.dtoh(cc.value, 256 * 4, np.int32), offsets=capi.dtoh(co.value, 257 * 4, np.int32),
                    batch_sizes=capi.dtoh(bs.value, 256 * 4, np.int32))

    def pgs_parts(self):
        """(work-groups per solver cell, constraints of its largest part) of the last solve"""
        parts, big = np.zeros(256, np.int32), np.zeros(256, np.int32)
        call("b3h_pipeline_get_pgs_parts", self.pl, ptr(parts), ptr(big))
        return parts, big

    def autotune(self, dt=1.0 / 60.0, steps=3):
        """(lanes per pair, PGS parts cap, ms per step) the autotuner keeps; the
        state is restored, so the simulation continues unchanged"""
        lanes, parts, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        call("b3h_pipeline_autotune", self.pl, ctypes.c_float(dt), int(steps), ctypes.byref(lanes),
             ctypes.byref(parts), ctypes.byref(ms))
        return lanes.value, parts.value, ms.value

    def static0_index(self):
        v = ctypes.c_int()
        call("b3h_np_get_static0_index", self.np_, ctypes.byref(v))
        return v.value

    def enable_timing(self, on=True):
        call("b3h_pipeline_enable_timing", self.pl, int(on))

    def stage_times(self):
        out = np.zeros(capi.NUM_STAGES, np.float32)
        call("b3h_pipeline_get_stage_times", self.pl, ptr(out))
        return out

    def track_residual(self, on=True):
        call("b3h_pipeline_set_track_residual", self.pl, int(on))

    def residuals(self):
        k = self.solver_params().m_numIterations
        out = np.zeros(2 * k, np.float32)
        call("b3h_pipeline_get_residuals", self.pl, ptr(out), len(out))
        return out

    def residuals_l2(self):
        k = self.solver_params().m_numIterations
        out = np.zeros(2 * k, np.float32)
        call("b3h_pipeline_get_residuals_l2", self.pl, ptr(out), len(out))
        return out

    # -- warm start (mode 0 cold as the reference, 1 previous frame, 2 learned)
    def set_warm_start(self, mode, net=None):
        if net is not None:
            call("b3h_pipeline_set_warm_start_model", self.pl, net.h)
        sp = self.solver_params()
        sp.m_warmStartMode = int(mode)
        self.set_solver(sp)

    def warm_debug(self, n):
        """(initial impulses [n][6], matched previous impulses [n][8], feature rows [4n][24] or None)"""
        w, pr, x = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        call("b3h_pipeline_get_warm_start_debug", self.pl, ctypes.byref(w), ctypes.byref(pr), ctypes.byref(x))
        warm = capi.dtoh(w.value, n * 24, np.float32).reshape(n, 6) if w.value else None
        prev = capi.dtoh(pr.value, n * 32, np.float32).reshape(n, 8) if pr.value else None
        feats = capi.dtoh(x.value, n * 4 * 96, np.float32).reshape(4 * n, 24) if x.value else None
        return warm, prev, feats

    def clear_warm_history(self):
        call("b3h_pipeline_clear_warm_start_history", self.pl)

    # -- batched scenes (body ranges; pairs across scenes are dropped)
    def set_scenes(self, offsets):
        off = np.ascontiguousarray(offsets, np.int32)
        self.num_scenes = len(off) - 1 if len(off) else 0
        call("b3h_pipeline_set_scenes", self.pl, self.num_scenes, ptr(off) if self.num_scenes else None)

    def scene_contact_counts(self):
        out = np.zeros(self.num_scenes, np.int32)
        call("b3h_pipeline_get_scene_contact_counts", self.pl, ptr(out), self.num_scenes)
        return out

    # -- state image (bodies, inertias, gravity, solver parameters, warm-start history)
    def save_state(self):
        n = ctypes.c_size_t()
        call("b3h_pipeline_save_state", self.pl, None, 0, ctypes.byref(n))
        buf = np.zeros(n.value, np.uint8)
        call("b3h_pipeline_save_state", self.pl, ptr(buf), len(buf), ctypes.byref(n))
        return buf[:n.value].tobytes()

    def load_state(self, image):
        buf = np.frombuffer(image, np.uint8).copy()
        call("b3h_pipeline_load_state", self.pl, ptr(buf), len(buf))

    def set_contact_log(self, x_dev, t_dev, capacity_rows):
        call("b3h_pipeline_set_contact_log", self.pl, ctypes.c_void_p(x_dev) if x_dev else None,
             ctypes.c_void_p(t_dev) if t_dev else None, int(capacity_rows))

    def contact_log_count(self):
        n = ctypes.c_int()
        call("b3h_pipeline_get_contact_log_count", self.pl, ctypes.byref(n))
        return n.value

    def overflow(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        call("b3h_pipeline_get_overflow", self.pl, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    # -- joints (createPoint2PointConstraint / createFixedConstraint)
    def add_point2point(self, a, b, pivot_a, pivot_b, breaking=1e30):
        uid = ctypes.c_int(-1)
        call("b3h_pipeline_create_point2point_constraint", self.pl, int(a), int(b),
             ptr(np.ascontiguousarray(pivot_a, np.float32)), ptr(np.ascontiguousarray(pivot_b, np.float32)),
             ctypes.c_float(breaking), ctypes.byref(uid))
        return uid.value

    def add_fixed(self, a, b, pivot_a, pivot_b, rel_target=(0.0, 0.0, 0.0, 1.0), breaking=1e30):
        uid = ctypes.c_int(-1)
        call("b3h_pipeline_create_fixed_constraint", self.pl, int(a), int(b),
             ptr(np.ascontiguousarray(pivot_a, np.float32)), ptr(np.ascontiguousarray(pivot_b, np.float32)),
             ptr(np.ascontiguousarray(rel_target, np.float32)), ctypes.c_float(breaking), ctypes.byref(uid))
        return uid.value

    def remove_constraint(self, uid):
        call("b3h_pipeline_remove_constraint_by_uid", self.pl, int(uid))

    def constraints(self):
        """the pipeline's constraint list, with the device's enabled flags"""
        n = ctypes.c_int()
        call("b3h_pipeline_get_constraints", self.pl, None, 0, ctypes.byref(n))
        out = np.zeros(n.value, layout.CONSTRAINT_JOINT)
        if n.value:
            call("b3h_pipeline_get_constraints", self.pl, ptr(out), n.value, ctypes.byref(n))
        return out

    def aabbs(self):
        """the broadphase's AABBs (getAllAabbsGPU), as the last step left them"""
        d, n = ctypes.c_void_p(), ctypes.c_int()
        call("b3h_bp_get_all_aabbs_gpu", self.bp, ctypes.byref(d), ctypes.byref(n))
        call("b3h_synchronize")
        return capi.dtoh(d.value, 32 * n.value, np.uint8).copy().view(layout.AABB)

    def cast_rays(self, rays, hits):
        """castRays on host arrays (layout.RAYINFO, layout.RAYHIT); returns the
        updated hit records"""
        rays = np.ascontiguousarray(rays)
        out = np.ascontiguousarray(hits).copy()
        call("b3h_pipeline_cast_rays", self.pl, ptr(rays), ptr(out), len(rays))
        return out

    def close(self):
        if self.pl:
            call("b3h_pipeline_destroy", self.pl)
            self.pl = ctypes.c_void_p()
        if self.bp:
            call("b3h_bp_destroy", self.bp)
            self.bp = ctypes.c_void_p()
        if self.np_:
            call("b3h_np_destroy", self.np_)
            self.np_ = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
