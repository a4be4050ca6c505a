"""Warm start on the device (k_warm.hip + the solver's warm pass) against the
oracle: the previous-frame match and the feature rows against
oracle/warm_oracle.py, the warm pass against the scalar solver's warm start
(bit for bit), a zero network against the previous-frame mode (bit for bit),
and the fixed point: warm and cold solves converge to the same velocities."""
import ctypes

import numpy as np
import pytest

import pyoracle as po
import warm_oracle as wo
from bullet3_amd import capi, layout, scenes
from bullet3_amd.world import World
from helpers import OracleWorld, bits_equal, first_diff

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
VEL_UNIT = np.float32(np.float32(9.8) * np.float32(DT))


@pytest.fixture(scope="module", autouse=True)
def need_device():
    if capi.device_count() < 1:
        pytest.skip("no HIP device")


def pile(seed=5):
    return scenes.convex_pile(6, 4, 6, spacing=1.0, n_hull_shapes=8, drop=0.0, seed=seed)


def cubes():
    return scenes.cube_grid(4, 4, 4, spacing=1.0, drop=0.0, seed=1)


@pytest.fixture(params=["cubes", "pile"])
def scene(request):
    return cubes() if request.param == "cubes" else pile()


def contacts_then_solve(w):
    """the step's collision stages, then the solve alone (no integration);
    returns the bodies the solve started from and the contacts"""
    capi.call("b3h_pipeline_setup_gpu_aabbs_full", w.pl)
    capi.call("b3h_bp_calculate_overlapping_pairs", w.bp, w.cfg.m_maxBroadphasePairs)
    pairs = w.pairs()
    d = ctypes.c_void_p()
    capi.call("b3h_bp_get_overlapping_pair_buffer", w.bp, ctypes.byref(d))
    capi.call("b3h_np_compute_contacts", w.np_, d, len(pairs), None, 0)
    contacts = w.contacts()
    b = w.bodies()
    capi.call("b3h_pipeline_solve_contacts", w.pl)
    return b, contacts


def solver_dt(w):
    return w.solver_params().m_solverDt


def test_previous_frame_match_and_warm_pass(scene):
    w = World(scene)
    w.set_warm_start(1)
    for _ in range(20):
        w.step(DT)
    prev_cons = w.solver_debug()["constraints"]
    b, contacts = contacts_then_solve(w)
    dbg = w.solver_debug()
    n = len(dbg["constraints"])
    warm, prev, _ = w.warm_debug(n)
    exp_prev = wo.match(prev_cons, dbg["constraints"])
    assert prev[:, 6].sum() > 0.5 * n          # most pairs persist
    assert np.array_equal(prev[:, 6:], exp_prev[:, 6:])
    assert np.allclose(prev[:, :6], exp_prev[:, :6], rtol=1e-6, atol=1e-9), \
        np.argwhere(~np.isclose(prev[:, :6], exp_prev[:, :6], rtol=1e-6, atol=1e-9))[:5]
    assert np.array_equal(warm, prev[:, :6])
    # the warm pass and the sweeps: the scalar solver from the same initial impulses
    ow = OracleWorld(w)
    cpu = po.solve_contacts(b, ow.inertias, contacts, ow.solver, static_idx=w.static0_index(),
                            initial_impulses=warm)
    assert bits_equal(dbg["constraints"], cpu["constraints"].view(layout.CONSTRAINT)), \
        first_diff(dbg["constraints"], cpu["constraints"].view(layout.CONSTRAINT))
    assert bits_equal(w.bodies(), b), first_diff(w.bodies(), b)
    w.close()


def test_feature_rows_and_targets(scene):
    w = World(scene)
    w.set_warm_start(1)
    for _ in range(15):
        w.step(DT)
    prev_cons = w.solver_debug()["constraints"]
    from bullet3_amd.learned import DeviceArray
    cap = 4 * w.cfg.m_maxContactCapacity
    lx, lt = DeviceArray((cap, 24)), DeviceArray((cap, 3))
    w.set_contact_log(lx.ptr, lt.ptr, cap)
    b, contacts = contacts_then_solve(w)
    cons = w.solver_debug()["constraints"]
    n = len(cons)
    _, prev, feats = w.warm_debug(n)
    exp_prev = wo.match(prev_cons, cons)
    pm = wo.point_matched(prev_cons, cons)
    X = wo.features(cons, b, exp_prev, pm, VEL_UNIT)
    assert np.allclose(feats, X, rtol=2e-5, atol=1e-5), \
        np.argwhere(~np.isclose(feats, X, rtol=2e-5, atol=1e-5))[:8]
    rows = w.contact_log_count()
    npts = exp_prev[:, 7].astype(int)
    assert rows == npts.sum()
    valid = np.concatenate([np.arange(4 * k, 4 * k + npts[k]) for k in range(n)])
    logged_x = lx.numpy()[:rows]
    assert np.array_equal(logged_x, feats[valid])
    T = wo.targets(cons, prev, VEL_UNIT)
    assert np.allclose(lt.numpy()[:rows], T[valid], rtol=1e-5, atol=1e-6)
    w.set_contact_log(None, None, 0)
    lx.free()
    lt.free()
    w.close()


def test_zero_network_equals_previous_frame():
    from bullet3_amd.learned import WarmStartNet
    sc = pile()
    a, b = World(sc), World(sc)
    a.set_warm_start(1)
    net = WarmStartNet(24, 3, max_train_rows=64)
    net.set_params(np.zeros(net.num_params, np.float32))
    b.set_warm_start(2, net)
    for s in range(30):
        a.step(DT)
        b.step(DT)
    assert np.array_equal(a.bodies().view(np.float32), b.bodies().view(np.float32))
    a.close()
    b.close()
    net.close()


def test_graph_replay_with_learned_warm_start():
    from bullet3_amd.learned import WarmStartNet
    sc = pile()
    a, b = World(sc), World(sc)
    net = WarmStartNet(24, 3, max_train_rows=64, seed=3)
    a.set_warm_start(2, net)
    b.set_warm_start(2, net)
    for rep in range(3):
        a.step_graph(DT, 4)
        for _ in range(4):
            b.step(DT)
        a.sync()
        assert bits_equal(a.bodies(), b.bodies()), "replay %d" % rep
    a.close()
    b.close()
    net.close()


def test_warm_and_cold_share_the_fixed_point():
    """only the initial iterate of each phase changes (normal impulses before
    the normal sweeps, friction impulses before the friction sweeps): with many
    sweeps the warm and cold solves reach the same velocities, far closer than
    after a few sweeps.  Each warm solve starts from the impulse history the
    solve before it left."""
    sc = scenes.cube_grid(4, 2, 4, spacing=1.0, drop=0.0, seed=1)
    w = World(sc)
    w.set_warm_start(1)
    for _ in range(40):
        w.step(DT)
    state = w.bodies()
    out = {}
    for iters in (2, 64):
        for mode in (1, 0):
            w.set_bodies(state)
            sp = w.solver_params()
            sp.m_numIterations = iters
            sp.m_warmStartMode = mode
            w.set_solver(sp)
            contacts_then_solve(w)
            out[(mode, iters)] = w.bodies()
    v = {k: np.concatenate([o["linVel"][:, :3], o["angVel"][:, :3]], 1) for k, o in out.items()}
    far = np.max(np.abs(v[(1, 2)] - v[(0, 2)]))
    near = np.max(np.abs(v[(1, 64)] - v[(0, 64)]))
    print("warm-cold max velocity difference: 2 sweeps %.3e, 64 sweeps %.3e" % (far, near), flush=True)
    assert far > 0
    assert near < 0.1 * far
    w.close()


@pytest.mark.parametrize("bad", ["nan", "huge"])
def test_bad_network_outputs_stay_safe(bad):
    """a network whose outputs are NaN falls back, constraint by constraint, to
    the previous-frame guess (so stepping equals the previous-frame mode bit
    for bit); outputs of 1e30 give finite, bounded motion (the warm pass clamps
    normal impulses at zero and friction to the friction bound)"""
    from bullet3_amd.learned import WarmStartNet
    sc = pile()
    a, b = World(sc), World(sc)
    a.set_warm_start(1)
    net = WarmStartNet(24, 3, max_train_rows=64)
    prm = np.zeros(net.num_params, np.float32)
    prm[-3:] = np.nan if bad == "nan" else 1e30  # output biases
    net.set_params(prm)
    b.set_warm_start(2, net)
    for s in range(30):
        a.step(DT)
        b.step(DT)
    bb = b.bodies()
    if bad == "nan":
        assert np.array_equal(a.bodies().view(np.float32), bb.view(np.float32))
    else:
        for f in ("pos", "linVel", "angVel"):
            assert np.isfinite(bb[f]).all()
    a.close()
    b.close()
    net.close()
