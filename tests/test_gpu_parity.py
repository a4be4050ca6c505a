"""GPU parity: every stage of the HIP path against the oracle on identical state.

The library and the oracle evaluate the same float operations in the same
order (no FMA contraction, correctly rounded division and square root, the same
polynomial sine/cosine), so integer outputs (pairs, slots, cells, batches) and
float outputs (AABBs, contacts, constraints, velocities, poses) are compared bit
for bit.  All calls go through the C ABI (bullet3_amd/libb3hip.so).
"""
import ctypes

import numpy as np
import pytest

import pyoracle as po
from bullet3_amd import capi, layout, scenes
from bullet3_amd.world import World
from helpers import OracleWorld, bits_equal, first_diff, sort_pairs

pytestmark = pytest.mark.gpu


def small_pile(n=(6, 4, 6), seed=5, spacing=1.0, drop=0.0):
    return scenes.convex_pile(*n, spacing=spacing, n_hull_shapes=8, drop=drop, seed=seed)


def small_cubes(seed=1):
    return scenes.cube_grid(4, 4, 4, spacing=1.0, drop=0.0, seed=seed)


@pytest.fixture(params=["cubes", "pile"])
def scene(request):
    return small_cubes() if request.param == "cubes" else small_pile()


def run_steps(world, ow, steps, dt=1.0 / 60.0):
    for _ in range(steps):
        world.step(dt)
        ow.step(dt)


def test_world_aabbs_bit_exact(scene):
    w = World(scene)
    ow = OracleWorld(w)
    capi.call("b3h_pipeline_setup_gpu_aabbs_full", w.pl)
    gpu = w.world_aabbs()
    cpu = po.world_aabbs(ow.bodies, ow.t["collidables"], ow.t["localAabbs"]).view(layout.AABB)
    assert bits_equal(gpu, cpu)
    w.close()


def test_integrate_bit_exact(scene):
    w = World(scene)
    rng = np.random.default_rng(0)
    b = w.bodies()
    dyn = b["invMass"] > 0
    b["linVel"][dyn, :3] = rng.normal(size=(dyn.sum(), 3))
    b["angVel"][dyn, :3] = rng.normal(size=(dyn.sum(), 3)) * 5.0
    idx = np.where(dyn)[0][:3]
    b["angVel"][idx, :3] = 200.0  # exercise the angular clamp
    w.set_bodies(b)
    for dt in (1.0 / 60.0, 1.0 / 240.0):
        capi.call("b3h_pipeline_integrate", w.pl, ctypes.c_float(dt))
        po.integrate(b, dt, 0.99, (0.0, -9.8, 0.0))
        assert bits_equal(w.bodies(), b), first_diff(w.bodies(), b)
    w.close()


def grid_hash_order(aabbs, small, cell, dims=128):
    """(cell hash, proxy) of every small proxy, stably sorted by hash
    (gridBroadphase.cl:38-49 hash of the AABB centre; float32 arithmetic)"""
    a = aabbs[small]
    c = (a["min"][:, :3] + a["max"][:, :3]) * np.float32(0.5)
    inv = np.float32(1.0) / np.float32(cell)
    g = np.floor(c * inv).astype(np.int64) & (dims - 1)
    h = (g[:, 2] * dims * dims + g[:, 1] * dims + g[:, 0]).astype(np.uint32)
    order = np.argsort(h, kind="stable").astype(np.uint32)
    return h[order], order


@pytest.mark.parametrize("kind", ["grid", "sap"])
def test_pairs_bit_exact(scene, kind):
    w = World(scene, broadphase=kind)
    ow = OracleWorld(w)
    run_steps(w, ow, 3)
    capi.call("b3h_pipeline_setup_gpu_aabbs_full", w.pl)
    capi.call("b3h_bp_calculate_overlapping_pairs", w.bp, w.cfg.m_maxBroadphasePairs)
    gpu = w.pairs()
    aabbs = w.world_aabbs()
    if kind == "grid":
        cs = w.cell_size
        kp, vp, nn = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int()
        capi.call("b3h_bp_get_grid_sort", w.bp, ctypes.byref(kp), ctypes.byref(vp), ctypes.byref(nn))
        keys = capi.dtoh(kp.value, 4 * nn.value, np.uint32)
        vals = capi.dtoh(vp.value, 4 * nn.value, np.uint32)
        ek, ev = grid_hash_order(aabbs, ow.small, cs)
        assert np.array_equal(keys, ek), "hash keys: %s" % first_diff(keys, ek)
        assert np.array_equal(vals, ev), "proxy order: %s" % first_diff(vals, ev)
        cpu, _ = po.grid_pairs(aabbs, ow.small, ow.large, cell=(cs, cs, cs))
    else:
        axis = ctypes.c_int()
        capi.call("b3h_bp_get_sap_axis", w.bp, ctypes.byref(axis))
        cpu, _, _ = po.sap_pairs(aabbs, ow.small, ow.large, axis=axis.value)
    assert len(gpu) > 0
    assert np.array_equal(gpu, cpu)
    brute = po.brute_pairs(aabbs, ow.small, ow.large)
    assert np.array_equal(sort_pairs(gpu), sort_pairs(brute))
    w.close()


@pytest.mark.parametrize("lanes", [64, 32, 16])
def test_narrowphase_bit_exact(scene, lanes):
    """lanes: lanes cooperating on one pair (64 = one pair per wavefront; 32 and
    16 pack two and four pairs into a wavefront)"""
    w = World(scene)
    capi.call("b3h_np_set_pairs_per_group", w.np_, lanes)
    ow = OracleWorld(w)
    run_steps(w, ow, 2)
    capi.call("b3h_pipeline_setup_gpu_aabbs_full", w.pl)
    capi.call("b3h_bp_calculate_overlapping_pairs", w.bp, w.cfg.m_maxBroadphasePairs)
    pairs = w.pairs()
    d = ctypes.c_void_p()
    capi.call("b3h_bp_get_overlapping_pair_buffer", w.bp, ctypes.byref(d))
    capi.call("b3h_np_compute_contacts", w.np_, d, len(pairs), None, 0)
    gpu_c = w.contacts()
    gpu_pairs = capi.dtoh(d.value, len(pairs) * 16, np.int32).reshape(-1, 4)
    b = w.bodies()
    n, cpu_c, cpu_pairs, has_sep, sep = po.narrowphase(pairs, b, ow.t, capacity=w.cfg.m_maxContactCapacity,
                                                       want_sep=True)
    st, sn = ctypes.c_void_p(), ctypes.c_void_p()
    capi.call("b3h_np_get_sat_debug", w.np_, ctypes.byref(st), ctypes.byref(sn))
    g_state = capi.dtoh(st.value, len(pairs) * 4, np.int32)
    g_sep = capi.dtoh(sn.value, len(pairs) * 16, np.float32).reshape(-1, 4)
    assert np.array_equal(g_state, has_sep)
    assert bits_equal(g_sep, sep)
    assert n == len(gpu_c) and n > 0
    assert np.array_equal(gpu_pairs, cpu_pairs)
    assert bits_equal(gpu_c, cpu_c.view(layout.CONTACT))
    w.close()


@pytest.mark.parametrize("lanes", [64, 32])
def test_narrowphase_shapes_changed_on_device(lanes):
    """Bodies whose collidable is rewritten in device memory only (the host
    mirror, and with it the per-body face-record ranges, still has the old
    shapes): the SAT projects those bodies' faces and fixed directions itself,
    the rest come from the records, and the results stay the oracle's."""
    w = World(small_pile(n=(6, 4, 6), seed=9))
    capi.call("b3h_np_set_pairs_per_group", w.np_, lanes)
    ow = OracleWorld(w)
    run_steps(w, ow, 2)
    b = w.bodies()
    dyn = np.nonzero(b["invMass"] > 0)[0]
    b2 = b.copy()
    b2["collidableIdx"][dyn] = np.roll(b["collidableIdx"][dyn], 1)
    assert np.count_nonzero(b2["collidableIdx"] != b["collidableIdx"]) > len(dyn) // 2
    w.set_bodies(b2)
    capi.call("b3h_pipeline_setup_gpu_aabbs_full", w.pl)
    capi.call("b3h_bp_calculate_overlapping_pairs", w.bp, w.cfg.m_maxBroadphasePairs)
    pairs = w.pairs()
    assert len(pairs) > 50
    d = ctypes.c_void_p()
    capi.call("b3h_bp_get_overlapping_pair_buffer", w.bp, ctypes.byref(d))
    capi.call("b3h_np_compute_contacts", w.np_, d, len(pairs), None, 0)
    gpu_c = w.contacts()
    st, sn = ctypes.c_void_p(), ctypes.c_void_p()
    capi.call("b3h_np_get_sat_debug", w.np_, ctypes.byref(st), ctypes.byref(sn))
    g_state = capi.dtoh(st.value, len(pairs) * 4, np.int32)
    g_sep = capi.dtoh(sn.value, len(pairs) * 16, np.float32).reshape(-1, 4)
    n, cpu_c, _, has_sep, sep = po.narrowphase(pairs, b2, ow.t, capacity=w.cfg.m_maxContactCapacity, want_sep=True)
    assert np.array_equal(g_state, has_sep)
    assert bits_equal(g_sep, sep)
    assert n == len(gpu_c) and n > 0
    assert bits_equal(gpu_c, cpu_c.view(layout.CONTACT)), first_diff(gpu_c, cpu_c.view(layout.CONTACT))
    w.close()


@pytest.mark.parametrize("cache", [-1, 0, 3])
def test_solver_bit_exact(scene, cache):
    """cache: bodies per cell kept in LDS (-1: the LDS budget; 0: every cell on
    the global walk; 3: mixed, small cells cached, larger ones global)"""
    w = World(scene)
    sp = w.solver_params()
    sp.m_numIterations = 10
    w.set_solver(sp)
    capi.call("b3h_pipeline_set_pgs_body_cache", w.pl, cache)
    ow = OracleWorld(w)
    run_steps(w, ow, 2)
    capi.call("b3h_pipeline_setup_gpu_aabbs_full", w.pl)
    capi.call("b3h_bp_calculate_overlapping_pairs", w.bp, w.cfg.m_maxBroadphasePairs)
    pairs = w.pairs()
    d = ctypes.c_void_p()
    capi.call("b3h_bp_get_overlapping_pair_buffer", w.bp, ctypes.byref(d))
    capi.call("b3h_np_compute_contacts", w.np_, d, len(pairs), None, 0)
    contacts = w.contacts()
    b = w.bodies()
    capi.call("b3h_pipeline_solve_contacts", w.pl)
    dbg = w.solver_debug()
    cpu = po.solve_contacts(b, ow.inertias, contacts, ow.solver, static_idx=w.static0_index())
    assert np.array_equal(dbg["counts"], cpu["counts"].astype(np.int32))
    assert np.array_equal(dbg["offsets"][:256], cpu["offsets"].astype(np.int32))
    assert np.array_equal(dbg["batch_sizes"], cpu["batch_sizes"])
    assert bits_equal(dbg["sorted_contacts"], cpu["sorted_contacts"].view(layout.CONTACT))
    assert bits_equal(dbg["constraints"], cpu["constraints"].view(layout.CONSTRAINT)), \
        first_diff(dbg["constraints"], cpu["constraints"].view(layout.CONSTRAINT))
    assert bits_equal(w.bodies(), b), first_diff(w.bodies(), b)
    w.close()


@pytest.mark.parametrize("parts,cache", [(1, -1), (8, -1), (8, 40), (4, 12)])
def test_solver_parts_bit_exact(parts, cache):
    """3 m solver cells wrap every 24 m in x and z and 12 m in y, so the cells
    of a 26 m wide pile hold two copies in space, which share no body: with
    up to 8 parts each such cell is solved by several work-groups at once,
    with 1 part by one.  Either way the sweep must equal the oracle's
    sequential one.  (Cells must stay wider than a body: cells of one parity
    class are one cell apart and are solved at once, as in the reference.)
    cache caps the bodies one work-group keeps in LDS: parts own part-local
    velocity slots, so a cell whose parts fit the cap is solved split while
    a cell with a larger part takes the global walk."""
    sc = small_pile(n=(26, 2, 26), spacing=1.0)
    w = World(sc)
    sp = w.solver_params()
    sp.m_numIterations = 10
    sp.m_batchCellSize = 3.0
    w.set_solver(sp)
    capi.call("b3h_pipeline_set_pgs_max_parts", w.pl, parts)
    capi.call("b3h_pipeline_set_pgs_body_cache", w.pl, cache)
    ow = OracleWorld(w)
    run_steps(w, ow, 2)
    assert bits_equal(w.bodies(), ow.bodies), first_diff(w.bodies(), ow.bodies)
    capi.call("b3h_pipeline_setup_gpu_aabbs_full", w.pl)
    capi.call("b3h_bp_calculate_overlapping_pairs", w.bp, w.cfg.m_maxBroadphasePairs)
    pairs = w.pairs()
    d = ctypes.c_void_p()
    capi.call("b3h_bp_get_overlapping_pair_buffer", w.bp, ctypes.byref(d))
    capi.call("b3h_np_compute_contacts", w.np_, d, len(pairs), None, 0)
    contacts = w.contacts()
    b = w.bodies()
    capi.call("b3h_pipeline_solve_contacts", w.pl)
    split, largest = w.pgs_parts()
    print("parts cap %d: cells split %d, parts max %d" % (parts, int(np.sum(split > 1)), int(split.max())), flush=True)
    if cache < 0 or cache == 40:
        assert (split.max() > 1) == (parts > 1)
    dbg = w.solver_debug()
    cpu = po.solve_contacts(b, ow.inertias, contacts, ow.solver, static_idx=w.static0_index())
    assert bits_equal(dbg["constraints"], cpu["constraints"].view(layout.CONSTRAINT)), \
        first_diff(dbg["constraints"], cpu["constraints"].view(layout.CONSTRAINT))
    assert bits_equal(w.bodies(), b), first_diff(w.bodies(), b)
    w.close()


@pytest.mark.parametrize("kind", ["grid", "sap"])
def test_steps_bit_exact(scene, kind):
    w = World(scene, broadphase=kind)
    sp = w.solver_params()
    sp.m_numIterations = 10
    sp.m_frictionCoeff = 0.5
    w.set_solver(sp)
    ow = OracleWorld(w, broadphase=0 if kind == "grid" else 1)
    for s in range(40):
        w.step()
        ow.step()
        if s % 10 == 9:
            gb = w.bodies()
            assert bits_equal(gb, ow.bodies), "step %d" % s
    assert ow.w.numContacts > 0
    w.close()


def test_steps_split_cells_bit_exact():
    """stepping (fused integration in the last friction sweep) with cells
    solved as several parts, against the oracle"""
    sc = small_pile(n=(26, 2, 26), spacing=1.0, drop=0.5)
    w = World(sc)
    sp = w.solver_params()
    sp.m_numIterations = 8
    sp.m_batchCellSize = 3.0
    w.set_solver(sp)
    ow = OracleWorld(w)
    split = 0
    for s in range(30):
        w.step()
        ow.step()
        split = max(split, int(w.pgs_parts()[0].max()))
        if s % 10 == 9:
            assert bits_equal(w.bodies(), ow.bodies), "step %d" % s
    assert split > 1 and ow.w.numContacts > 0
    w.close()


def test_graph_replay_matches_eager():
    sc = small_pile()
    a = World(sc)
    b = World(sc)
    for rep in range(4):
        a.step_graph(1.0 / 60.0, 5)
        for _ in range(5):
            b.step()
        a.sync()
        ba, bb = a.bodies(), b.bodies()
        assert bits_equal(ba, bb), "replay %d" % rep
    a.close()
    b.close()


def test_overflow_is_reported_and_clamped():
    sc = small_pile()
    w = World(sc, max_bodies=sc.num_bodies + 10, pairs_per_body=1)  # tiny pair capacity
    for _ in range(3):
        w.step()
    po_, co = w.overflow()
    assert po_ == 1
    assert len(w.pairs()) == w.cfg.m_maxBroadphasePairs
    w.close()
