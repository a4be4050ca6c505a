"""ctypes binding of the b3hip C ABI (include/b3hip.h).

The shared library bullet3_amd/libb3hip.so is built in-tree by
`make -C bullet3_amd/csrc` (or __graft_entry__.build()).  Loading fails loudly
when it is missing: there is no CPU fallback behind this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libb3hip.so")
_LIB = None

B3H_OK = 0
NUM_STAGES = 9
STAGE_NAMES = ("aabb", "broadphase", "narrowphase_mpr", "narrowphase_sat", "narrowphase_compact", "solver_setup",
               "pgs_normal", "pgs_friction", "integrate")
BROADPHASE_GRID = 0
BROADPHASE_SAP = 1
BROADPHASE_PLBVH = 2
BROADPHASE_KINDS = {"grid": BROADPHASE_GRID, "sap": BROADPHASE_SAP, "plbvh": BROADPHASE_PLBVH}


class B3hError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "m_maxConvexBodies", "m_maxConvexShapes", "m_maxBroadphasePairs", "m_maxContactCapacity",
        "m_compoundPairCapacity", "m_maxVerticesPerFace", "m_maxFacesPerShape", "m_maxConvexVertices",
        "m_maxConvexIndices", "m_maxConvexUniqueEdges", "m_maxCompoundChildShapes", "m_maxTriConvexPairCapacity")]


class SolverParams(ctypes.Structure):
    _fields_ = [("m_solverDt", ctypes.c_float), ("m_numIterations", ctypes.c_int),
                ("m_frictionCoeff", ctypes.c_float), ("m_positionDrift", ctypes.c_float),
                ("m_positionConstraintCoeff", ctypes.c_float), ("m_batchCellSize", ctypes.c_float),
                ("m_angularDamping", ctypes.c_float), ("m_warmStartMode", ctypes.c_int)]


class ConvexUtility(ctypes.Structure):
    """b3hConvexUtility: a precomputed polyhedron (pointers to host arrays)."""
    _fields_ = [("vertices", ctypes.c_void_p), ("numVertices", ctypes.c_int), ("numFaces", ctypes.c_int),
                ("facePlanes", ctypes.c_void_p), ("faceIndexCounts", ctypes.c_void_p), ("indices", ctypes.c_void_p),
                ("uniqueEdges", ctypes.c_void_p), ("numUniqueEdges", ctypes.c_int), ("pad", ctypes.c_int),
                # b3hFloat4 is 16-byte aligned: mC starts at byte 64
                ("pad16", ctypes.c_int * 2),
                ("mC", ctypes.c_float * 4), ("mE", ctypes.c_float * 4), ("m_extents", ctypes.c_float * 4),
                ("m_radius", ctypes.c_float), ("localCenterOut", ctypes.c_void_p)]


assert ctypes.sizeof(ConvexUtility) == 128 and ConvexUtility.mC.offset == 64 and ConvexUtility.localCenterOut.offset == 120


_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_pi = ctypes.POINTER(ctypes.c_int)
_pvp = ctypes.POINTER(ctypes.c_void_p)
_sz = ctypes.c_size_t

_SIGS = {
    "b3h_last_error": ([], ctypes.c_char_p),
    "b3h_version": ([], ctypes.c_char_p),
    "b3h_config_default": ([ctypes.POINTER(Config)], _i),
    "b3h_solver_params_default": ([ctypes.POINTER(SolverParams)], _i),
    "b3h_np_create": ([ctypes.POINTER(Config), _vp, _pvp], _i),
    "b3h_np_destroy": ([_vp], _i),
    "b3h_np_register_convex_hull_shape": ([_vp, _vp, _i, _i, _vp, _pi], _i),
    "b3h_np_register_sphere_shape": ([_vp, _f, _pi], _i),
    "b3h_np_register_plane_shape": ([_vp, _vp, _f, _pi], _i),
    "b3h_np_register_compound_shape": ([_vp, _vp, _i, _pi], _i),
    "b3h_np_get_compound_tables": ([_vp, _vp, _vp, _vp, _vp, _vp], _i),
    "b3h_compute_compound_tables": ([_vp, _i, _vp, _vp, _vp, _vp, _i, _pi, _vp, _i, _pi], _i),
    "b3h_np_register_convex_polyhedron": ([_vp, ctypes.POINTER(ConvexUtility), _pi], _i),
    "b3h_np_load_shape_tables": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "b3h_np_get_mpr_debug": ([_vp, _pvp], _i),
    "b3h_bp_calculate_overlapping_pairs_host": ([_vp, _i], _i),
    "b3h_bp_get_all_aabbs_gpu": ([_vp, _pvp, _pi], _i),
    "b3h_bp_get_all_aabbs_cpu": ([_vp, _vp, _i, _pi], _i),
    "b3h_bp_get_small_aabb_indices_gpu": ([_vp, _pvp, _pi], _i),
    "b3h_bp_get_large_aabb_indices_gpu": ([_vp, _pvp, _pi], _i),
    "b3h_pipeline_register_convex_polyhedron": ([_vp, ctypes.POINTER(ConvexUtility), _pi], _i),
    "b3h_compute_body_registration": ([_vp, _vp, _f, _vp, _vp, _vp, _vp, _vp], _i),
    "b3h_np_register_rigid_body": ([_vp, _i, _f, _vp, _vp, _vp, _vp, _i, _pi], _i),
    "b3h_np_write_all_bodies_to_gpu": ([_vp], _i),
    "b3h_np_reset": ([_vp], _i),
    "b3h_np_readback_all_bodies_to_cpu": ([_vp, _vp], _i),
    "b3h_np_set_object_transform_cpu": ([_vp, _vp, _vp, _i], _i),
    "b3h_np_set_object_velocity_cpu": ([_vp, _vp, _vp, _i], _i),
    "b3h_np_get_object_transform_from_cpu": ([_vp, _vp, _vp, _i], _i),
    "b3h_np_compute_contacts": ([_vp, _vp, _i, _vp, _i], _i),
    "b3h_np_get_bodies_gpu": ([_vp, _pvp, _pi], _i),
    "b3h_np_get_body_inertias_gpu": ([_vp, _pvp, _pi], _i),
    "b3h_np_get_collidables_gpu": ([_vp, _pvp, _pi], _i),
    "b3h_np_get_contacts_gpu": ([_vp, _pvp, _pi], _i),
    "b3h_np_get_static0_index": ([_vp, _pi], _i),
    "b3h_np_get_num_rigid_bodies": ([_vp, _pi], _i),
    "b3h_np_get_local_space_aabb": ([_vp, _i, _vp], _i),
    "b3h_np_get_shape_tables": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "b3h_np_get_inertias_cpu": ([_vp, _vp, _i], _i),
    "b3h_np_set_pairs_per_group": ([_vp, _i], _i),
    "b3h_np_get_sat_debug": ([_vp, _pvp, _pvp], _i),
    "b3h_bp_create": ([_i, ctypes.POINTER(Config), _vp, _pvp], _i),
    "b3h_bp_destroy": ([_vp], _i),
    "b3h_bp_set_grid_params": ([_vp, _vp, _vp, _i], _i),
    "b3h_bp_create_proxy": ([_vp, _vp, _vp, _i, _i, _i], _i),
    "b3h_bp_create_large_proxy": ([_vp, _vp, _vp, _i, _i, _i], _i),
    "b3h_bp_write_aabbs_to_gpu": ([_vp], _i),
    "b3h_bp_calculate_overlapping_pairs": ([_vp, _i], _i),
    "b3h_bp_get_aabb_buffer_ws": ([_vp, _pvp, _pi], _i),
    "b3h_bp_get_num_overlap": ([_vp, _pi], _i),
    "b3h_bp_get_overlapping_pair_buffer": ([_vp, _pvp], _i),
    "b3h_bp_get_total_found": ([_vp, _pi], _i),
    "b3h_bp_get_sap_axis": ([_vp, _pi], _i),
    "b3h_bp_get_bvh_info": ([_vp, _pi, _pi], _i),
    "b3h_pipeline_create_point2point_constraint": ([_vp, _i, _i, _vp, _vp, _f, _pi], _i),
    "b3h_pipeline_create_fixed_constraint": ([_vp, _i, _i, _vp, _vp, _vp, _f, _pi], _i),
    "b3h_pipeline_remove_constraint_by_uid": ([_vp, _i], _i),
    "b3h_pipeline_copy_constraints_to_host": ([_vp], _i),
    "b3h_pipeline_get_constraints": ([_vp, _vp, _i, _pi], _i),
    "b3h_pipeline_cast_rays": ([_vp, _vp, _vp, _i], _i),
    "b3h_pipeline_cast_rays_gpu": ([_vp, _vp, _vp, _i], _i),
    "b3h_pipeline_get_ray_bvh_info": ([_vp, _pi, _pi], _i),
    "b3h_bp_get_proxy_maps": ([_vp, _vp, _vp, _vp], _i),
    "b3h_bp_create_proxies": ([_vp, _i, _vp, _vp, _vp], _i),
    "b3h_bp_get_grid_sort": ([_vp, _pvp, _pvp, _pi], _i),
    "b3h_pipeline_create": ([_vp, _vp, ctypes.POINTER(Config), _vp, _pvp], _i),
    "b3h_pipeline_destroy": ([_vp], _i),
    "b3h_pipeline_step_simulation": ([_vp, _f], _i),
    "b3h_pipeline_integrate": ([_vp, _f], _i),
    "b3h_pipeline_setup_gpu_aabbs_full": ([_vp], _i),
    "b3h_pipeline_solve_contacts": ([_vp], _i),
    "b3h_pipeline_register_physics_instance": ([_vp, _f, _vp, _vp, _i, _i, _i, _pi], _i),
    "b3h_pipeline_write_all_instances_to_gpu": ([_vp], _i),
    "b3h_pipeline_set_gravity": ([_vp, _vp], _i),
    "b3h_pipeline_reset": ([_vp], _i),
    "b3h_pipeline_get_body_buffer": ([_vp, _pvp], _i),
    "b3h_pipeline_get_num_bodies": ([_vp, _pi], _i),
    "b3h_pipeline_set_solver_params": ([_vp, ctypes.POINTER(SolverParams)], _i),
    "b3h_pipeline_get_solver_params": ([_vp, ctypes.POINTER(SolverParams)], _i),
    "b3h_pipeline_get_solver_debug": ([_vp, _pvp, _pvp, _pvp, _pvp, _pvp, _pi], _i),
    "b3h_pipeline_set_pgs_body_cache": ([_vp, _i], _i),
    "b3h_pipeline_get_pgs_levels": ([_vp, _vp], _i),
    "b3h_pipeline_get_pgs_parts": ([_vp, _vp, _vp], _i),
    "b3h_pipeline_set_pgs_max_parts": ([_vp, _i], _i),
    "b3h_pipeline_autotune": ([_vp, _f, _i, _pi, _pi, _vp], _i),
    "b3h_pipeline_set_track_residual": ([_vp, _i], _i),
    "b3h_pipeline_get_residuals": ([_vp, _vp, _i], _i),
    "b3h_pipeline_get_residuals_l2": ([_vp, _vp, _i], _i),
    "b3h_pipeline_set_warm_start_impulses": ([_vp, _vp, _i], _i),
    "b3h_pipeline_set_warm_start_model": ([_vp, _vp], _i),
    "b3h_pipeline_clear_warm_start_history": ([_vp], _i),
    "b3h_pipeline_save_state": ([_vp, _vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)], _i),
    "b3h_pipeline_load_state": ([_vp, _vp, ctypes.c_size_t], _i),
    "b3h_pipeline_set_scenes": ([_vp, _i, _vp], _i),
    "b3h_pipeline_get_scene_contact_counts": ([_vp, _vp, _i], _i),
    "b3h_pipeline_get_warm_start_debug": ([_vp, _pvp, _pvp, _pvp], _i),
    "b3h_pipeline_set_contact_log": ([_vp, _vp, _vp, _i], _i),
    "b3h_pipeline_get_contact_log_count": ([_vp, _pi], _i),
    "b3h_pipeline_enable_timing": ([_vp, _i], _i),
    "b3h_pipeline_get_stage_times": ([_vp, _vp], _i),
    "b3h_pipeline_get_overflow": ([_vp, _pi, _pi], _i),
    "b3h_pipeline_step_graph": ([_vp, _f, _i], _i),
    "b3h_mlp_create": ([_i, _i, _i, ctypes.c_uint, _vp, _pvp], _i),
    "b3h_mlp_destroy": ([_vp], _i),
    "b3h_mlp_get_dims": ([_vp, _pi, _pi, _pi], _i),
    "b3h_mlp_num_params": ([_vp, _pi], _i),
    "b3h_mlp_set_params": ([_vp, _vp], _i),
    "b3h_mlp_get_params": ([_vp, _vp], _i),
    "b3h_mlp_set_input_normalization": ([_vp, _vp, _vp], _i),
    "b3h_mlp_get_input_normalization": ([_vp, _vp, _vp], _i),
    "b3h_mlp_forward": ([_vp, _vp, _i, _i, _vp, _i], _i),
    "b3h_mlp_loss_gradients": ([_vp, _vp, _i, _vp, _i, _i, _vp, _vp], _i),
    "b3h_mlp_train": ([_vp, _vp, _i, _vp, _i, _i, _i, _i, _f, ctypes.c_uint, _vp], _i),
    "b3h_radix_sort32": ([_vp, _vp, _i, _i], _i),
    "b3h_radix_sort32_noalloc": ([_vp, _vp, _i, _i], _i),
    "b3h_prefix_scan": ([_vp, _vp, _i, _pi], _i),
    "b3h_compute_polyhedral_features": ([_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "b3h_device_count": ([_pi], _i),
    "b3h_malloc": ([_pvp, _sz], _i),
    "b3h_free": ([_vp], _i),
    "b3h_memcpy_htod": ([_vp, _vp, _sz], _i),
    "b3h_memcpy_dtoh": ([_vp, _vp, _sz], _i),
    "b3h_memcpy_dtod": ([_vp, _vp, _sz], _i),
    "b3h_synchronize": ([], _i),
}


def exported_symbols():
    return list(_SIGS.keys())


def lib():
    """The loaded C-ABI library (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise B3hError("bullet3_amd/libb3hip.so is not built; run __graft_entry__.build() "
                           "or make -C bullet3_amd/csrc")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def check(rc, what=""):
    if rc != B3H_OK:
        msg = lib().b3h_last_error().decode(errors="replace")
        raise B3hError("%s failed (%d): %s" % (what, rc, msg))


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    check(rc, name)


def ptr(a):
    if a is None:
        return None
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    # data_as keeps a reference to the array, so a temporary argument stays
    # alive for the duration of the foreign call
    return a.ctypes.data_as(ctypes.c_void_p)


def default_config(max_bodies=None, pairs_per_body=16):
    c = Config()
    call("b3h_config_default", ctypes.byref(c))
    if max_bodies is not None:
        c.m_maxConvexBodies = int(max_bodies)
        c.m_maxConvexShapes = int(max_bodies)
        c.m_maxBroadphasePairs = int(pairs_per_body * max_bodies)
        c.m_maxContactCapacity = c.m_maxBroadphasePairs
    return c


def default_solver_params():
    p = SolverParams()
    call("b3h_solver_params_default", ctypes.byref(p))
    return p


def device_count():
    n = ctypes.c_int(0)
    call("b3h_device_count", ctypes.byref(n))
    return n.value


def dtoh(dev_ptr, nbytes, dtype=np.uint8):
    out = np.empty(nbytes, np.uint8)
    if nbytes:
        call("b3h_memcpy_dtoh", ptr(out), ctypes.c_void_p(dev_ptr), nbytes)
    return out.view(dtype)


def htod(dev_ptr, arr):
    arr = np.ascontiguousarray(arr)
    if arr.nbytes:
        call("b3h_memcpy_htod", ctypes.c_void_p(dev_ptr), ptr(arr), arr.nbytes)


class DeviceBuffer:
    """Raw device allocation through the C ABI (for tests and benchmarks)."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        call("b3h_malloc", ctypes.byref(p), int(nbytes))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def free(self):
        if self.ptr:
            call("b3h_free", ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def polyhedral_features(points, cap=4096):
    """Hull features as registerConvexHullShape stores them (host-only call)."""
    pts = np.ascontiguousarray(points, np.float32)
    counts = np.zeros(4, np.int32)
    v = np.zeros((cap, 4), np.float32)
    p = np.zeros((cap, 4), np.float32)
    fc = np.zeros(cap, np.int32)
    idx = np.zeros(cap, np.int32)
    e = np.zeros((cap, 4), np.float32)
    call("b3h_compute_polyhedral_features", ptr(pts), 12, len(pts), cap, ptr(counts), ptr(v), ptr(p), ptr(fc),
         ptr(idx), ptr(e))
    nv, nf, ni, ne = (int(x) for x in counts)
    faces = []
    off = 0
    for f in range(nf):
        faces.append(idx[off:off + fc[f]].copy())
        off += fc[f]
    return dict(vertices=v[:nv, :3].copy(), planes=p[:nf].copy(), faces=faces, edges=e[:ne, :3].copy())


def register_convex_polyhedron(np_handle, tables, pipeline=None):
    """registerConvexHullShape(b3ConvexUtility*) with precomputed tables
    {vertices (n,3|4), planes (f,4), faces [index arrays], edges (e,3|4)} and
    optional mC / mE / extents (4 floats) / radius; returns the collidable
    index (through the pipeline's registerConvexPolyhedron when `pipeline` is
    given)."""
    def f4(a):
        a = np.asarray(a, np.float32).reshape(len(a), -1)
        out = np.zeros((len(a), 4), np.float32)
        out[:, :a.shape[1]] = a[:, :4]
        return np.ascontiguousarray(out)

    v = f4(tables["vertices"])
    p = np.ascontiguousarray(tables["planes"], np.float32).reshape(-1, 4)
    fc = np.array([len(f) for f in tables["faces"]], np.int32)
    idx = np.ascontiguousarray(np.concatenate([np.asarray(f, np.int32) for f in tables["faces"]])
                               if len(fc) else np.zeros(0, np.int32), np.int32)
    e = f4(tables["edges"]) if len(tables["edges"]) else np.zeros((0, 4), np.float32)
    center = np.zeros(4, np.float32)
    u = ConvexUtility()
    u.vertices, u.numVertices = v.ctypes.data, len(v)
    u.numFaces, u.facePlanes, u.faceIndexCounts = len(fc), p.ctypes.data, fc.ctypes.data
    u.indices = idx.ctypes.data
    u.uniqueEdges, u.numUniqueEdges = (e.ctypes.data if len(e) else None), len(e)
    for name in ("mC", "mE", "m_extents"):
        val = tables.get(name)
        if val is not None:
            getattr(u, name)[:] = [float(x) for x in val]
    u.m_radius = float(tables.get("radius", 0.0))
    u.localCenterOut = center.ctypes.data
    out = ctypes.c_int(-1)
    if pipeline is not None:
        call("b3h_pipeline_register_convex_polyhedron", pipeline, ctypes.byref(u), ctypes.byref(out))
    else:
        call("b3h_np_register_convex_polyhedron", np_handle, ctypes.byref(u), ctypes.byref(out))
    tables["localCenter"] = center
    return out.value


def radix_sort32(keys, values, sort_bits=32, entry="b3h_radix_sort32"):
    """Stable (key, value) sort on the device by the low sort_bits key bits
    (entry: b3h_radix_sort32 or its allocation-free variant)."""
    k = np.ascontiguousarray(keys, np.uint32)
    v = np.ascontiguousarray(values, np.uint32)
    n = len(k)
    dk, dv = DeviceBuffer(4 * max(n, 1)), DeviceBuffer(4 * max(n, 1))
    htod(dk.ptr, k)
    htod(dv.ptr, v)
    call(entry, ctypes.c_void_p(dk.ptr), ctypes.c_void_p(dv.ptr), n, sort_bits)
    return dtoh(dk.ptr, 4 * n, np.uint32).copy(), dtoh(dv.ptr, 4 * n, np.uint32).copy()


def prefix_scan(values):
    """Exclusive prefix sum on the device; returns (scan of n+1 ints, total)."""
    a = np.ascontiguousarray(values, np.int32)
    n = len(a)
    ds, dd = DeviceBuffer(4 * max(n, 1)), DeviceBuffer(4 * (n + 1))
    htod(ds.ptr, a)
    total = ctypes.c_int()
    call("b3h_prefix_scan", ctypes.c_void_p(ds.ptr), ctypes.c_void_p(dd.ptr), n, ctypes.byref(total))
    return dtoh(dd.ptr, 4 * (n + 1), np.int32).copy(), total.value
