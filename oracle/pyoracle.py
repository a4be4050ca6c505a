"""ORACLE -- test infrastructure only.

ctypes binding of oracle/liboracle.so (the scalar CPU restatement of the
Bullet3 GPU rigid-body step).  Imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py only; the product never loads it.
Arrays are numpy arrays holding the reference byte layouts (structured or raw).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _LIB = ctypes.CDLL(path)
    return _LIB


def ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    # data_as keeps a reference to the array (temporaries stay alive)
    return a.ctypes.data_as(ctypes.c_void_p)


class SolverParams(ctypes.Structure):
    _fields_ = [("m_solverDt", ctypes.c_float), ("m_numIterations", ctypes.c_int),
                ("m_frictionCoeff", ctypes.c_float), ("m_positionDrift", ctypes.c_float),
                ("m_positionConstraintCoeff", ctypes.c_float), ("m_batchCellSize", ctypes.c_float),
                ("m_angularDamping", ctypes.c_float), ("m_warmStartMode", ctypes.c_int)]


def set_threads(n):
    """threads the oracle may use (results do not depend on it)"""
    lib().b3o_set_threads(int(n))


def solver_params(dt=1.0 / 60.0, iterations=4, friction=0.7, drift=0.005, coeff=0.2, cell=6.0, damping=0.99):
    return SolverParams(dt, iterations, friction, drift, coeff, cell, damping, 0)


class CompoundTables(ctypes.Structure):
    _fields_ = [("children", ctypes.c_void_p), ("infos", ctypes.c_void_p), ("nodes", ctypes.c_void_p),
                ("subtrees", ctypes.c_void_p), ("capacity", ctypes.c_int)]


def compound_tables(tables, capacity):
    """CompoundTables over the arrays of tables (keys children, infos, nodes,
    subtrees; the arrays must outlive the struct)"""
    t = CompoundTables()
    if tables is None or capacity <= 0 or len(tables.get("children", ())) == 0:
        return t
    t.children = tables["children"].ctypes.data
    t.infos = tables["infos"].ctypes.data
    t.nodes = tables["nodes"].ctypes.data
    t.subtrees = tables["subtrees"].ctypes.data
    t.capacity = capacity
    return t


class World(ctypes.Structure):
    _fields_ = [("numBodies", ctypes.c_int), ("bodies", ctypes.c_void_p), ("inertias", ctypes.c_void_p),
                ("collidables", ctypes.c_void_p), ("convexes", ctypes.c_void_p), ("vertices", ctypes.c_void_p),
                ("uniqueEdges", ctypes.c_void_p), ("faces", ctypes.c_void_p), ("indices", ctypes.c_void_p),
                ("localAabbs", ctypes.c_void_p), ("smallMap", ctypes.c_void_p), ("nSmall", ctypes.c_int),
                ("largeMap", ctypes.c_void_p), ("nLarge", ctypes.c_int), ("broadphase", ctypes.c_int),
                ("cellSize", ctypes.c_float * 3), ("gridDims", ctypes.c_int * 3), ("maxBodiesPerCell", ctypes.c_int),
                ("maxPairs", ctypes.c_int), ("maxContacts", ctypes.c_int), ("solver", SolverParams),
                ("gravity", ctypes.c_float * 3), ("staticIdx", ctypes.c_int),
                ("worldAabbs", ctypes.c_void_p), ("pairs", ctypes.c_void_p), ("contacts", ctypes.c_void_p),
                ("numPairs", ctypes.c_int), ("numContacts", ctypes.c_int), ("residuals", ctypes.c_void_p),
                ("constraints", ctypes.c_void_p), ("numConstraints", ctypes.c_int), ("compound", CompoundTables)]


def world_aabbs(bodies, collidables, local_aabbs):
    n = len(bodies)
    out = np.zeros((n, 32), np.uint8)
    lib().b3o_compute_world_aabbs(n, ptr(bodies), ptr(collidables), ptr(local_aabbs), ptr(out))
    return out


def integrate(bodies, dt, damping=0.99, gravity=(0.0, -9.8, 0.0)):
    g = np.array(gravity, np.float32)
    lib().b3o_integrate(len(bodies), ptr(bodies), ctypes.c_float(dt), ctypes.c_float(damping), ptr(g))


def grid_pairs(aabbs, small, large, cell=(3.0, 3.0, 3.0), dims=(128, 128, 128), max_per_cell=256, max_pairs=None):
    small = np.ascontiguousarray(small, np.int32)
    large = np.ascontiguousarray(large, np.int32)
    if max_pairs is None:
        max_pairs = 64 * max(1, len(small)) + 16
    out = np.zeros((max_pairs, 4), np.int32)
    total = ctypes.c_int(0)
    c = np.array(cell, np.float32)
    d = np.array(dims, np.int32)
    n = lib().b3o_grid_pairs(ptr(aabbs), ptr(small), len(small), ptr(large), len(large), ptr(c), ptr(d),
                             max_per_cell, ptr(out), max_pairs, ctypes.byref(total))
    return out[:n].copy(), total.value


def sap_pairs(aabbs, small, large, axis=-1, max_pairs=None):
    small = np.ascontiguousarray(small, np.int32)
    large = np.ascontiguousarray(large, np.int32)
    if max_pairs is None:
        max_pairs = 64 * max(1, len(small)) + 16
    out = np.zeros((max_pairs, 4), np.int32)
    total = ctypes.c_int(0)
    ax = ctypes.c_int(0)
    n = lib().b3o_sap_pairs(ptr(aabbs), ptr(small), len(small), ptr(large), len(large), axis, ptr(out), max_pairs,
                            ctypes.byref(total), ctypes.byref(ax))
    return out[:n].copy(), total.value, ax.value


def plbvh_pairs(aabbs, small, large, max_pairs=None):
    """(pairs in the device's order, total found, root node)"""
    small = np.ascontiguousarray(small, np.int32)
    large = np.ascontiguousarray(large, np.int32)
    if max_pairs is None:
        max_pairs = 64 * max(1, len(small)) + 16
    out = np.zeros((max_pairs, 4), np.int32)
    total = ctypes.c_int(0)
    root = ctypes.c_int(0)
    n = lib().b3o_plbvh_pairs(ptr(aabbs), ptr(small), len(small), ptr(large), len(large), ptr(out), max_pairs,
                              ctypes.byref(total), ctypes.byref(root))
    return out[:n].copy(), total.value, root.value


def world_aabbs(bodies, collidables, local_aabbs):
    """broadphase AABBs of the bodies (b3ComputeWorldAabb, margin 0)"""
    from bullet3_amd import layout
    out = np.zeros(len(bodies), layout.AABB)
    lib().b3o_compute_world_aabbs(len(bodies), ptr(bodies), ptr(collidables), ptr(local_aabbs), ptr(out))
    return out


def cast_rays(aabbs, small, large, bodies, tables, rays, hits):
    """nearest hit of each ray (hits: RAYHIT records, updated copy returned)"""
    small = np.ascontiguousarray(small, np.int32)
    large = np.ascontiguousarray(large, np.int32)
    rays = np.ascontiguousarray(rays)
    out = np.ascontiguousarray(hits).copy()
    lib().b3o_cast_rays(ptr(aabbs), ptr(small), len(small), ptr(large), len(large), ptr(bodies),
                        ptr(tables["collidables"]), ptr(tables["convexes"]), ptr(tables["faces"]), ptr(rays), ptr(out),
                        len(rays))
    return out


def solve_joints(bodies, inertias, constraints, time_step=1.0 / 60.0, iterations=4):
    """one joint solve in place (bodies' velocities, constraints' flags);
    returns the number of batches"""
    return lib().b3o_solve_joints(len(bodies), ptr(bodies), ptr(inertias), ptr(constraints), len(constraints),
                                  ctypes.c_float(time_step), int(iterations))


def brute_pairs(aabbs, small, large, max_pairs=None):
    small = np.ascontiguousarray(small, np.int32)
    large = np.ascontiguousarray(large, np.int32)
    if max_pairs is None:
        max_pairs = 64 * max(1, len(small)) + 16
    out = np.zeros((max_pairs, 4), np.int32)
    n = lib().b3o_brute_pairs(ptr(aabbs), ptr(small), len(small), ptr(large), len(large), ptr(out), max_pairs)
    return out[:min(n, max_pairs)].copy()


def mpr_contacts(pairs, bodies, collidables, convexes, vertices, capacity=None):
    pairs = np.ascontiguousarray(pairs, np.int32).copy()
    if capacity is None:
        capacity = len(pairs)
    contacts = np.zeros((max(capacity, 1), 112), np.uint8)
    n = lib().b3o_mpr_contacts(ptr(pairs), len(pairs), ptr(bodies), ptr(collidables), ptr(convexes), ptr(vertices),
                               ptr(contacts), capacity)
    return n, contacts[:min(n, capacity)].copy(), pairs


def narrowphase(pairs, bodies, shapes, capacity=None, want_sep=False, compound=None, compound_capacity=0,
                want_child_pairs=False):
    """compound: dict of the compound tables (children, infos, nodes,
    subtrees) with compound_capacity child pairs kept; want_child_pairs adds
    the child-pair list to the result"""
    pairs = np.ascontiguousarray(pairs, np.int32).copy()
    if capacity is None:
        capacity = max(1, len(pairs))
    contacts = np.zeros((capacity, 112), np.uint8)
    has_sep = np.zeros(max(1, len(pairs)), np.int32) if want_sep else None
    sep = np.zeros((max(1, len(pairs)), 4), np.float32) if want_sep else None
    ct = compound_tables(compound, compound_capacity)
    child = np.zeros((max(1, compound_capacity), 4), np.int32)
    nchild = ctypes.c_int(0)
    n = lib().b3o_narrowphase_ex(ptr(pairs), len(pairs), ptr(bodies), ptr(shapes["collidables"]),
                                 ptr(shapes["convexes"]), ptr(shapes["vertices"]), ptr(shapes["uniqueEdges"]),
                                 ptr(shapes["faces"]), ptr(shapes["indices"]), ptr(contacts), capacity,
                                 ptr(has_sep), ptr(sep), ctypes.byref(ct), ptr(child), ctypes.byref(nchild))
    res = (n, contacts[:min(n, capacity)].copy(), pairs)
    if want_sep:
        res = res + (has_sep[:len(pairs)], sep[:len(pairs)])
    if want_child_pairs:
        res = res + (child[:min(nchild.value, compound_capacity)].copy(), nchild.value)
    return res


def solve_contacts(bodies, inertias, contacts, params, static_idx=0, initial_impulses=None):
    n = len(contacts)
    sorted_c = np.zeros((max(n, 1), 112), np.uint8)
    cons = np.zeros((max(n, 1), 176), np.uint8)
    counts = np.zeros(256, np.uint32)
    offsets = np.zeros(256, np.uint32)
    batches = np.zeros(256, np.int32)
    resid = np.zeros(2 * params.m_numIterations, np.float32)
    init = None if initial_impulses is None else np.ascontiguousarray(initial_impulses, np.float32)
    lib().b3o_solve_contacts(len(bodies), ptr(bodies), ptr(inertias), n, ptr(np.ascontiguousarray(contacts)),
                             ctypes.byref(params), static_idx, ptr(sorted_c), ptr(cons), ptr(counts), ptr(offsets),
                             ptr(batches), ptr(init), ptr(resid))
    return dict(sorted_contacts=sorted_c[:n], constraints=cons[:n], counts=counts, offsets=offsets,
                batch_sizes=batches, residuals=resid)
