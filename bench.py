"""Benchmark: Bullet3 GPU rigid-body step on one MI355X.

Headline workload: the single large pile -- 131072 dynamic convex bodies (one
third unit boxes, 16- and 32-point hulls), a static ground box, uniform-grid
broadphase with cell 2*sqrt(3) m, 10 PGS iterations, dt 1/60.  A "step" is one
full stepSimulation: world AABBs, broadphase, narrowphase, solver set-up, 2x10
PGS sweeps, integration.

Phases:
  1. settle: --settle graph-launched steps (default 300) that let the lattice
     collapse into a resting pile; the autotuner runs half-way through.  This
     count is fixed and reported as config.settle_steps, so --warmup does not
     change which state is timed.
  2. warm-up: --warmup more graph-launched steps (untimed).
  3. timed: exactly --steps graph-launched steps (eager with --eager),
     bracketed by device synchronisation.  value = body-steps/s over them.

Also in the one JSON line:
  roofline      the kernel with the largest share of the step: its
                algorithmic HBM bytes per launch (records read once, written
                once; kernel_bytes below) over its average launch duration,
                measured with HIP events on the pipeline's stream during eager
                steps right after the timed region.  traffic: measured HBM
                bytes per launch of that kernel from profiles/traffic.json,
                used only when that file was measured at this run's state
                (bodies, graph-launched steps before the stage timing, lanes
                per pair: 620 for the defaults), else null.
  stages_ms     per-stage HIP-event times of the same eager steps.
  warm_start    PGS iterations to reach a residual, learned warm start
                against plain PGS (cold, as the reference, and previous-frame).
  cpu_baseline  the oracle (the same algorithm in scalar C++) stepped from the
                GPU's settled state: one thread (value) and all host threads.
  configs       the other workloads at their sizes: 1000 cubes, 1,048,576
                ray casts into the pile, 1,048,576 AABBs (grid, SAP and BVH
                pairs/s), 32768 bodies of mixed shapes (compounds, spheres, a
                plane) with 2048 joints, 4096 batched scenes.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DT = 1.0 / 60.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def kernel_bytes(n, p, c):
    """algorithmic HBM bytes per launch of the kernels that can dominate a step:
    narrowphase MPR: pair 16 + two poses 2x32 read, direction/point/depth 48
    written per pair; SAT/clip: pair 16 + two poses 64 + MPR result 48 read and
    a flag 4 written per pair, contact 112 written per contact; a PGS launch
    (one of the 8 cell-parity launches of a sweep): constraint row data 176
    read + 16 impulses written per contact, velocities 32 read + 32 written,
    world inverse inertia 48 and position 16 per body"""
    sweep = 192 * c + 128 * n
    return {
        "narrowphase_mpr": 128 * p,
        "narrowphase_sat": 132 * p + 112 * c,
        "pgs_normal": sweep / 8.0,
        "pgs_friction": sweep / 8.0,
    }


def step_bytes(n, p, c, k):
    return 304 * n + 96 * p + 668 * c + 2 * k * (192 * c + 128 * n)


def host_threads():
    """threads for the all-core CPU figure: the cores this process may use,
    capped at 16 (the GPU box's CPU share for one GPU)"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def make_pile_world(args):
    from bullet3_amd import capi, scenes
    from bullet3_amd.world import World
    sc = scenes.convex_pile(args.nx, args.ny, args.nz)
    w = World(sc, broadphase="grid")
    sp = w.solver_params()
    sp.m_numIterations = args.iters
    if args.friction is not None:
        sp.m_frictionCoeff = args.friction
    w.set_solver(sp)
    if args.lanes:
        capi.call("b3h_np_set_pairs_per_group", w.np_, args.lanes)
    return sc, w


def run_pile(args):
    from bullet3_amd import capi
    from bullet3_amd.capi import STAGE_NAMES
    t0 = time.time()
    sc, w = make_pile_world(args)
    n_dyn = int((sc.mass > 0).sum())
    setup_s = time.time() - t0
    log("pile: %d bodies registered in %.1f s; settling %d steps" % (n_dyn, setup_s, args.settle))
    tuned = None
    half = args.settle // 2
    if half:
        w.step_graph(DT, half)
    if not args.no_autotune:
        # narrowphase lanes per pair and PGS parts per cell picked by timing trial
        # steps from a saved state image, which is restored afterwards
        lanes, parts, tms = w.autotune(DT, steps=args.autotune_steps)
        if args.lanes:
            capi.call("b3h_np_set_pairs_per_group", w.np_, args.lanes)
            lanes = args.lanes
        tuned = {"lanes_per_pair": lanes, "pgs_parts": parts, "trial_ms": round(tms, 4)}
    if args.settle - half:
        w.step_graph(DT, args.settle - half)
    if args.warmup:
        w.step_graph(DT, args.warmup)
    w.sync()
    settled_contacts = len(w.contacts())
    log("settled: %d manifolds; timing %d steps" % (settled_contacts, args.steps))
    # timed region
    w.sync()
    t0 = time.perf_counter()
    if args.eager:
        for _ in range(args.steps):
            w.step(DT)
    else:
        w.step_graph(DT, args.steps)
    w.sync()
    sec = time.perf_counter() - t0
    value = n_dyn * args.steps / sec
    ms_per_step = 1000.0 * sec / args.steps
    # per-stage times from HIP events on the pipeline's stream (eager steps, same state)
    w.enable_timing(True)
    reps = 5
    acc = np.zeros(len(STAGE_NAMES))
    for _ in range(reps):
        w.step(DT)
        acc += w.stage_times()
    w.enable_timing(False)
    st = dict(zip(STAGE_NAMES, (acc / reps).tolist()))
    n_pairs = len(w.pairs())
    n_contacts = len(w.contacts())
    lanes_used = ctypes.c_int()
    launches = {"narrowphase_mpr": 1, "narrowphase_sat": 1, "pgs_normal": 8 * args.iters,
                "pgs_friction": 8 * args.iters}
    kb = kernel_bytes(sc.num_bodies, n_pairs, n_contacts)
    dom = max(launches, key=lambda k: st[k])
    avg_ms = st[dom] / launches[dom]
    achieved = kb[dom] / (avg_ms * 1e-3) / 1e9
    total_bytes = step_bytes(sc.num_bodies, n_pairs, n_contacts, args.iters)
    lanes_now = tuned["lanes_per_pair"] if tuned else (args.lanes or 64)
    traffic, traffic_note = measured_traffic(dom, n_dyn, args.settle + args.warmup + args.steps, lanes_now)
    out = {
        "metric": "body-steps/s at 128k convexes; PGS iters to residual 1e-4, learned vs plain",
        "value": round(value, 1),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "convex pile %d dynamic bodies (boxes, 16/32-point hulls) on a ground box, grid "
                               "broadphase, %d PGS iterations, dt 1/60, settled %d steps" % (n_dyn, args.iters,
                                                                                            args.settle),
                   "bodies": n_dyn, "settle_steps": args.settle, "pairs": n_pairs, "contacts": n_contacts,
                   "contacts_when_settled": settled_contacts, "pgs_iterations": args.iters,
                   "setup_s": round(setup_s, 2), "autotune": tuned, "launch": "eager" if args.eager else "hipGraph"},
        "stages_ms": {k: round(float(v), 4) for k, v in st.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_note": traffic_note, "launch_ms": round(avg_ms, 5), "bytes_per_launch": int(kb[dom]),
                     "step_algorithmic_bytes": int(total_bytes),
                     "step_frac": round(total_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
    }
    del lanes_used
    state = w.bodies()
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(w, args)
        if out["cpu_baseline"]["value"] > 0:
            out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    w.set_bodies(state)
    return out, w, sc


def measured_traffic(kernel, bodies, steps_before, lanes):
    """HBM bytes per launch from profiles/traffic.json when it was measured at
    this state (the pile after the same number of graph-launched steps) and
    launch shape, else (None, reason)"""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None, "no measurement"
    with open(path) as f:
        t = json.load(f)
    s = t.get("state", {})
    if s.get("bodies") != bodies or s.get("steps_before") != steps_before or s.get("lanes_per_pair") != lanes:
        return None, "profiles/traffic.json was measured at another state or lane count (%s)" % s
    k = t.get("kernels", {}).get(kernel)
    if not k:
        return None, "kernel not measured"
    return int(k["bytes_per_launch"]), "rocprofv3 FETCH_SIZE+WRITE_SIZE at this state, %s" % t.get("scale", "")


def cpu_baseline(w, args):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pyoracle as po
    from helpers import OracleWorld
    ow = OracleWorld(w)
    n_dyn = int((ow.bodies["invMass"] > 0).sum())
    po.set_threads(1)
    t0 = time.perf_counter()
    for _ in range(args.cpu_steps_1):
        ow.step(DT)
    sec1 = time.perf_counter() - t0
    nt = host_threads()
    po.set_threads(nt)
    t0 = time.perf_counter()
    for _ in range(args.cpu_steps):
        ow.step(DT)
    secn = time.perf_counter() - t0
    po.set_threads(1)
    return {"value": round(n_dyn * args.cpu_steps_1 / sec1, 1), "unit": "body-steps/s", "cores": 1, "kind": "port",
            "sample": "%d oracle steps on one thread from the GPU's settled pile (%.1f s)" % (args.cpu_steps_1, sec1),
            "all_cores": {"value": round(n_dyn * args.cpu_steps / secn, 1), "cores": nt,
                          "sample": "the next %d oracle steps on %d threads (%.1f s; same bytes as one thread)" % (
                              args.cpu_steps, nt, secn)}}


def learned_eval(w, args, extra_worlds=()):
    """PGS iterations to a residual, learned warm start against plain PGS.

    Training rows: every contact point of --learn-collect-steps steps of the
    pile (and --learn-collect-batched steps of each extra world, the batched
    scenes) solved by previous-frame warm-start PGS at --learn-collect-iters
    iterations (near-converged impulses as targets); the network trains on the
    device (--learn-train-steps Adam steps of 8192 rows).  Then for cold start
    (the reference), previous-frame and learned warm start and each iteration
    count, --learn-eval-steps steps from the pile's settled state (no impulse
    history at the first); the residual of a step is the relative L2 change of
    its last normal sweep, averaged over the steps after --learn-skip.
    Reported: iterations each mode needs to reach plain 20-iteration PGS's
    residual and 1e-4, and graph-timed ms/step at those counts."""
    from bullet3_amd import warmstart as ws
    from bullet3_amd.learned import DeviceArray, WarmStartNet
    state = w.bodies()
    sp0 = w.solver_params()
    t0 = time.time()
    cap = args.learn_capacity
    X = DeviceArray((cap, 24))
    T = DeviceArray((cap, 3))
    rows = ws.collect_into(w, X, T, 0, args.learn_collect_steps, dt=DT, mode=ws.PREVIOUS,
                           iterations=args.learn_collect_iters)
    pile_rows = rows
    for ew in extra_worlds:
        rows = ws.collect_into(ew, X, T, rows, args.learn_collect_batched, dt=DT, mode=ws.PREVIOUS,
                               iterations=args.learn_collect_iters)
    collect_s = time.time() - t0
    w.set_bodies(state)
    net = WarmStartNet(24, 3, max_train_rows=8192, seed=1)
    t0 = time.time()
    hist = ws.train(net, X, T, rows, args.learn_train_steps, batch=8192, lr=1e-3, lr_final=1e-4, normalize=True)
    train_s = time.time() - t0
    X.free()
    T.free()
    iters = [int(k) for k in args.learn_iters.split(",")]
    curves = {}
    for mode in (ws.COLD, ws.PREVIOUS, ws.LEARNED):
        curves[mode] = {}
        for k in iters:
            curves[mode][k], _ = ws.run_residuals(w, state, mode, k, args.learn_eval_steps, args.learn_skip, dt=DT,
                                                  net=net)
        log("warm start: mode %s done" % ws.MODE_NAMES[mode])
    name = ws.MODE_NAMES
    targets = {"plain_cold_20": curves[ws.COLD][20], "plain_warm_20": curves[ws.PREVIOUS][20], "1e-4": 1e-4}
    k_to = {t: {name[m]: ws.iterations_to(curves[m], r) for m in curves} for t, r in targets.items()}
    ms = {}

    def timed(mode, k):
        key = "%s_%d" % (name[mode], k)
        if key not in ms:
            ms[key] = round(ws.time_steps(w, state, mode, k, args.learn_time_steps, dt=DT), 4)
        return ms[key]

    speed = {}
    for t in ("plain_cold_20", "plain_warm_20"):
        base = ws.COLD if t == "plain_cold_20" else ws.PREVIOUS
        k_l = k_to[t]["learned"]
        if k_l:
            speed[t] = {"iteration_ratio": round(20.0 / k_l, 3),
                        "steps_per_s_ratio": round(timed(base, 20) / timed(ws.LEARNED, k_l), 3)}
        else:
            speed[t] = None
    for m in curves:
        k = k_to["1e-4"][name[m]]
        if k:
            timed(m, k)
    w.set_bodies(state)
    w.clear_warm_history()
    w.set_solver(sp0)
    net.close()
    return {
        "residual": "relative L2 change of the last normal sweep, mean over steps",
        "targets": {t: float("%.4e" % r) for t, r in targets.items()},
        "iterations_to": k_to,
        "learned_vs_plain_20": speed,
        "ms_per_step": ms,
        "curves": {name[m]: {str(k): float("%.4e" % v) for k, v in curves[m].items()} for m in curves},
        "training": {"rows": int(rows), "pile_rows": int(pile_rows), "batched_rows": int(rows - pile_rows),
                     "collect_steps": args.learn_collect_steps, "collect_batched_steps": args.learn_collect_batched,
                     "collect_iterations": args.learn_collect_iters, "collect_s": round(collect_s, 2),
                     "adam_steps": args.learn_train_steps, "batch": 8192, "train_s": round(train_s, 2),
                     "loss_first_100": float(np.mean(hist[:100])) if len(hist) else None,
                     "loss_last_100": float(np.mean(hist[-100:])) if len(hist) else None},
    }


def sub_cubes(args):
    """the reference's CPU scene: 10x10x10 unit cubes dropped onto a ground box,
    friction 0.5, 10 PGS iterations, 200 steps; GPU graph-launched against the
    one-thread oracle over the same 200 steps from the same start"""
    from bullet3_amd import scenes
    from bullet3_amd.world import World
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pyoracle as po
    from helpers import OracleWorld
    sc = scenes.cube_grid(10, 10, 10)
    w = World(sc)
    sp = w.solver_params()
    sp.m_numIterations = 10
    sp.m_frictionCoeff = 0.5
    w.set_solver(sp)
    ow = OracleWorld(w)
    steps = 200
    w.step_graph(DT, 1)
    w.sync()
    w.set_bodies(ow.bodies)
    w.sync()
    t0 = time.perf_counter()
    w.step_graph(DT, steps)
    w.sync()
    sec = time.perf_counter() - t0
    po.set_threads(1)
    t0 = time.perf_counter()
    for _ in range(steps):
        ow.step(DT)
    csec = time.perf_counter() - t0
    same = w.bodies().tobytes() == ow.bodies.tobytes()
    w.close()
    return {"workload": "10x10x10 unit cubes on a 200x1x200 m ground, friction 0.5, 10 PGS iterations, 200 steps",
            "bodies": 1000, "steps": steps, "value": round(1000 * steps / sec, 1), "unit": "body-steps/s",
            "ms_per_step": round(1000 * sec / steps, 4),
            "cpu_port_one_thread": round(1000 * steps / csec, 1), "gpu_equals_oracle_after_200_steps": same}


def sub_shapes(args, n=32768, joints=2048, settle=100, steps=100):
    """the shapes beyond hulls in one scene: n dynamic bodies cycling through a
    unit box, a 16-point hull, three compounds (two small boxes; a hull, a box
    and a small box under child rotations; one box at the origin) and spheres,
    on a ground box and a static plane, with point-to-point joints between
    x-neighbours of the lattice; 10 PGS iterations.  settle graph-launched
    steps, then body-steps/s over steps more; stage times on 5 eager steps;
    the one-thread oracle (which restates the compound, primitive and joint
    paths as well) on 1 step from the same state"""
    from bullet3_amd import scenes
    from bullet3_amd.capi import STAGE_NAMES
    from bullet3_amd.world import World
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pyoracle as po
    from helpers import OracleWorld
    sc = scenes.compound_scene(n=n, seed=11, ground_half=(40.0, 0.5, 40.0))
    w = World(sc, broadphase="grid")
    sp = w.solver_params()
    sp.m_numIterations = 10
    w.set_solver(sp)
    # bodies 0, 1 are the ground box and the plane
    jl = scenes.lattice_joints(n, joints)
    for a, b, pa, pb in jl:
        w.add_point2point(a, b, pa, pb)
    made = len(jl)
    w.step_graph(DT, settle)
    w.sync()
    t0 = time.perf_counter()
    w.step_graph(DT, steps)
    w.sync()
    sec = time.perf_counter() - t0
    w.enable_timing(True)
    acc = np.zeros(len(STAGE_NAMES))
    for _ in range(5):
        w.step(DT)
        acc += w.stage_times()
    w.enable_timing(False)
    c = w.contacts()
    child = int(np.count_nonzero((c["childA"] >= 0) | (c["childB"] >= 0)))
    out = {"workload": "%d dynamic bodies (unit boxes, 16-point hulls, 3 compound shapes, spheres) on a ground box "
                       "and a plane, %d point-to-point joints, grid broadphase, 10 PGS iterations, settled %d steps"
                       % (n, made, settle),
           "bodies": n, "joints": made, "steps": steps, "value": round(n * steps / sec, 1), "unit": "body-steps/s",
           "ms_per_step": round(1000 * sec / steps, 4), "pairs": len(w.pairs()), "contacts": len(c),
           "compound_child_contacts": child,
           "stages_ms": {k: round(float(v) / 5, 4) for k, v in zip(STAGE_NAMES, acc)}}
    if not args.no_cpu:
        ow = OracleWorld(w)
        po.set_threads(1)
        t0 = time.perf_counter()
        ow.step(DT)
        csec = time.perf_counter() - t0
        out["cpu_port_one_thread"] = round(n / csec, 1)
    w.close()
    return out


def sub_aabbs(args):
    """broadphase stress: 1,048,576 AABBs (half-extents log-uniform 0.05-1 m,
    about 8 partners each), grid (cell 2 m), SAP and the linear BVH; pairs/s
    over repeated calculateOverlappingPairs; one-thread oracle grid timed once"""
    from bullet3_amd import capi, scenes
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    mn, mx, side = scenes.aabb_cloud(1 << 20)
    n = len(mn)
    res = {"workload": "1048576 AABBs, half-extents log-uniform 0.05-1.0 m, cube side %.1f m" % side, "aabbs": n}
    aabbs_host = None
    for kind, name in ((capi.BROADPHASE_GRID, "grid"), (capi.BROADPHASE_SAP, "sap"),
                       (capi.BROADPHASE_PLBVH, "plbvh")):
        cfg = capi.default_config(n + 16, 16)
        bp = ctypes.c_void_p()
        capi.call("b3h_bp_create", kind, ctypes.byref(cfg), None, ctypes.byref(bp))
        cell = np.array([2.0, 2.0, 2.0], np.float32)
        dims = np.array([128, 128, 128], np.int32)
        capi.call("b3h_bp_set_grid_params", bp, capi.ptr(cell), capi.ptr(dims), 256)
        capi.call("b3h_bp_create_proxies", bp, n, capi.ptr(np.ascontiguousarray(mn)),
                  capi.ptr(np.ascontiguousarray(mx)), None)
        capi.call("b3h_bp_write_aabbs_to_gpu", bp)
        for _ in range(2):
            capi.call("b3h_bp_calculate_overlapping_pairs", bp, cfg.m_maxBroadphasePairs)
        capi.call("b3h_synchronize")
        reps = 10
        t0 = time.perf_counter()
        for _ in range(reps):
            capi.call("b3h_bp_calculate_overlapping_pairs", bp, cfg.m_maxBroadphasePairs)
        capi.call("b3h_synchronize")
        sec = (time.perf_counter() - t0) / reps
        npairs = ctypes.c_int()
        capi.call("b3h_bp_get_num_overlap", bp, ctypes.byref(npairs))
        res[name] = {"pairs": npairs.value, "ms": round(1000 * sec, 4), "pairs_per_s": round(npairs.value / sec, 1)}
        if aabbs_host is None:
            dev = ctypes.c_void_p()
            cnt = ctypes.c_int()
            capi.call("b3h_bp_get_aabb_buffer_ws", bp, ctypes.byref(dev), ctypes.byref(cnt))
            aabbs_host = capi.dtoh(dev.value, 32 * cnt.value).copy()
        capi.call("b3h_bp_destroy", bp)
    from bullet3_amd import layout
    a = aabbs_host.view(layout.AABB)
    small = np.arange(n, dtype=np.int32)
    large = np.zeros(0, np.int32)
    t0 = time.perf_counter()
    cp, _ = po.grid_pairs(a, small, large, cell=(2.0, 2.0, 2.0), max_pairs=16 * n)
    csec = time.perf_counter() - t0
    res["cpu_port_grid_one_thread"] = {"pairs": len(cp), "ms": round(1000 * csec, 1),
                                       "pairs_per_s": round(len(cp) / csec, 1)}
    return res


def sub_rays(w, n_rays=1 << 20, seed=3):
    """ray casts into the settled pile (castRays on device arrays): n_rays
    segments from 45 m above the pile straight down to 5 m below the ground
    top, uniform over the pile's footprint; rays/s over repeated casts"""
    from bullet3_amd import capi, layout
    b = w.bodies()
    dyn = b["invMass"] > 0
    lo, hi = b["pos"][dyn, :3].min(0), b["pos"][dyn, :3].max(0)
    rng = np.random.default_rng(seed)
    r = np.zeros(n_rays, layout.RAYINFO)
    h = np.zeros(n_rays, layout.RAYHIT)
    xs = rng.uniform(lo[0], hi[0], n_rays).astype(np.float32)
    zs = rng.uniform(lo[2], hi[2], n_rays).astype(np.float32)
    r["from"][:, 0], r["from"][:, 1], r["from"][:, 2] = xs, 45.0, zs
    r["to"][:, 0], r["to"][:, 1], r["to"][:, 2] = xs, -5.0, zs
    h["hitBody"], h["hitResult2"] = -1, -1
    dr, dh, d0 = capi.DeviceBuffer(r.nbytes), capi.DeviceBuffer(h.nbytes), capi.DeviceBuffer(h.nbytes)
    capi.htod(dr.ptr, r)
    capi.htod(dh.ptr, h)
    capi.htod(d0.ptr, h)
    cast = lambda: capi.call("b3h_pipeline_cast_rays_gpu", w.pl, ctypes.c_void_p(dr.ptr), ctypes.c_void_p(dh.ptr),
                             n_rays)
    cast()
    w.sync()
    reps = 10
    sec = 0.0
    for _ in range(reps):
        capi.call("b3h_memcpy_dtod", ctypes.c_void_p(dh.ptr), ctypes.c_void_p(d0.ptr), h.nbytes)
        capi.call("b3h_synchronize")
        t0 = time.perf_counter()
        cast()
        w.sync()
        sec += time.perf_counter() - t0
    sec /= reps
    out = capi.dtoh(dh.ptr, h.nbytes, np.uint8).copy().view(layout.RAYHIT)
    hits = int((out["hitBody"] >= 0).sum())
    dr.free()
    dh.free()
    d0.free()
    return {"workload": "%d vertical rays over the settled pile's footprint, 50 m long" % n_rays, "rays": n_rays,
            "hits": hits, "ms": round(1000 * sec, 4), "rays_per_s": round(n_rays / sec, 1),
            "note": "host clock around each cast (BVH build and traversal), device synchronised on both sides"}


def sub_batched(args):
    """4096 independent scenes (16-512 bodies, piles and stacks, own grounds) in
    one pipeline: one launch per stage over all scenes, hipGraph-captured step,
    autotuned lanes per pair; 500 timed steps after 100 settle steps"""
    from bullet3_amd import scenes
    from bullet3_amd.world import World
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pyoracle as po
    from helpers import OracleWorld
    sc = scenes.batched_scenes(args.scenes)
    n_dyn = int((sc.mass > 0).sum())
    w = World(sc, broadphase="grid")
    sp = w.solver_params()
    sp.m_numIterations = args.iters
    w.set_solver(sp)
    w.set_scenes(sc.meta["scene_offsets"])
    w.step_graph(DT, 50)
    lanes, parts, _ = w.autotune(DT, steps=2)
    w.step_graph(DT, 50)
    w.sync()
    steps = args.batched_steps
    t0 = time.perf_counter()
    w.step_graph(DT, steps)
    w.sync()
    sec = time.perf_counter() - t0
    res = {"workload": "%d batched scenes, %d dynamic bodies, grid broadphase, %d PGS iterations" % (
        args.scenes, n_dyn, args.iters), "bodies": n_dyn, "steps": steps, "value": round(n_dyn * steps / sec, 1),
        "unit": "body-steps/s", "ms_per_step": round(1000 * sec / steps, 4), "pairs": len(w.pairs()),
        "contacts": len(w.contacts()), "autotune": {"lanes_per_pair": lanes, "pgs_parts": parts}}
    if not args.no_cpu:
        ow = OracleWorld(w)
        nt = host_threads()
        po.set_threads(nt)
        t0 = time.perf_counter()
        ow.step(DT)
        csec = time.perf_counter() - t0
        po.set_threads(1)
        res["cpu_port_all_cores"] = {"value": round(n_dyn / csec, 1), "cores": nt, "steps": 1}
    return res, w


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--settle", type=int, default=300,
                    help="graph-launched steps that let the pile settle before --warmup (fixed: the timed state)")
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--friction", type=float, default=None)
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=32)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--lanes", type=int, default=32,
                    help="narrowphase lanes per pair (64, 32, 16; 0 lets the autotuner pick).  Fixed at 32 by "
                         "default: at the pile the SAT stage takes 2.03 / 2.06 / 2.44 ms at 32 / 16 / 64 lanes, and "
                         "a fixed choice keeps the measured traffic (profiles/traffic.json, 32 lanes) applicable")
    ap.add_argument("--scenes", type=int, default=4096, help="scenes of the batched workload")
    ap.add_argument("--batched-steps", type=int, default=500)
    ap.add_argument("--cpu-steps", type=int, default=10, help="all-core oracle steps")
    ap.add_argument("--cpu-steps-1", type=int, default=2, help="one-thread oracle steps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start evaluation")
    ap.add_argument("--no-configs", action="store_true", help="skip the cubes / AABB / batched sub-results")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--autotune-steps", type=int, default=3)
    ap.add_argument("--eager", action="store_true",
                    help="launch every step's kernels one by one instead of replaying a captured HIP graph")
    ap.add_argument("--learn-collect-steps", type=int, default=40)
    ap.add_argument("--learn-collect-batched", type=int, default=10)
    ap.add_argument("--learn-capacity", type=int, default=64_000_000)
    ap.add_argument("--learn-train-steps", type=int, default=12000)
    ap.add_argument("--learn-collect-iters", type=int, default=60)
    ap.add_argument("--learn-eval-steps", type=int, default=16)
    ap.add_argument("--learn-skip", type=int, default=6)
    ap.add_argument("--learn-time-steps", type=int, default=30)
    ap.add_argument("--learn-iters", default="2,3,4,5,6,8,10,12,14,16,18,20,30,40,60,80,100,150,200,300,500")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path here (the step would shard by scene "
              "or by spatial slab); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from bullet3_amd import capi
    if capi.device_count() < 1:
        raise SystemExit("bench.py needs a HIP device")
    out, w, _ = run_pile(args)
    extra = []
    if not args.no_configs:
        log("config: cubes")
        out["configs"] = {"cubes_1000": sub_cubes(args)}
        log("config: rays")
        out["configs"]["rays_1m"] = sub_rays(w)
        log("config: 1M AABBs")
        out["configs"]["aabbs_1m"] = sub_aabbs(args)
        log("config: shapes")
        out["configs"]["shapes_32k"] = sub_shapes(args)
        log("config: batched scenes")
        out["configs"]["batched_%d" % args.scenes], bw = sub_batched(args)
        extra.append(bw)
    if not args.no_learned:
        log("learned warm start")
        out["warm_start"] = learned_eval(w, args, extra)
    for ew in extra:
        ew.close()
    w.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
