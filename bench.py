"""Benchmark: Bullet3 GPU rigid-body step on one MI355X.

Workload (default): the single large pile -- 131072 dynamic convex bodies (one
third unit boxes, 16- and 32-point hulls), a static ground box, uniform-grid
broadphase with cell 2*sqrt(3) m, 10 PGS iterations, dt 1/60.  Warmup steps let
the lattice collapse into a pile (and capture the step into a HIP graph); then
exactly --steps steps are timed.  A "step" is one full stepSimulation: world
AABBs, broadphase, narrowphase, solver set-up, 2x10 PGS sweeps, integration.

Prints one JSON line (metric body-steps/s).  roofline: the kernel with the
largest share of the step (the narrowphase SAT/clipping kernel or its MPR
kernel, or the normal or the friction PGS sweep launches), its algorithmic HBM
bytes per launch (records read once and written once, kernel_bytes below) over
its average launch duration, measured with HIP events on the pipeline's stream
during eager steps taken right after the timed region.  traffic: measured HBM
bytes per launch of that kernel from profiles/traffic.json (rocprof FETCH_SIZE
and WRITE_SIZE runs scaled by a copy of known size).  cpu_baseline: the scalar oracle
(same algorithm, one core) stepped from the GPU's post-warmup state for ~10 s.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def workload_name(args, n_dyn):
    if args.config == "pile":
        return ("convex pile %d dynamic bodies (boxes, 16/32-point hulls), grid broadphase, %d PGS iterations, "
                "dt 1/60" % (n_dyn, args.iters))
    if args.config == "batched":
        return ("%d batched scenes, %d dynamic bodies (piles of boxes and 8-24-point hulls, box stacks, 16-512 "
                "bodies each, own grounds), grid broadphase, %d PGS iterations, dt 1/60" % (args.scenes, n_dyn,
                                                                                            args.iters))
    return "10x10x10 unit cubes, grid broadphase, %d PGS iterations, dt 1/60" % args.iters


def kernel_bytes(n, p, c):
    """algorithmic HBM bytes per launch of the kernels that can dominate a step
    (records read once, written once: body 80, pair 16, contact 112,
    constraint 176 with 16 of impulses written, velocities 32 read + 32 written,
    world inverse inertia 48, position 16)"""
    sweep = 192 * c + 128 * n
    return {
        "narrowphase_mpr": 128 * p,               # pair + 2 poses read, direction/point/depth written
        "narrowphase_sat": 128 * p + 112 * c,     # pair + 2 poses + MPR result read, contact written
        "pgs_normal": sweep / 8.0,                # one of the 8 cell-parity launches of a sweep
        "pgs_friction": sweep / 8.0,
    }


def step_bytes(n, p, c, k):
    return 304 * n + 96 * p + 668 * c + 2 * k * (192 * c + 128 * n)


def run(args):
    from bullet3_amd import capi, scenes
    from bullet3_amd.world import World

    if capi.device_count() < 1:
        raise SystemExit("bench.py needs a HIP device")
    if args.config == "pile":
        sc = scenes.convex_pile(args.nx, args.ny, args.nz)
    elif args.config == "batched":
        sc = scenes.batched_scenes(args.scenes)
    else:
        sc = scenes.cube_grid(10, 10, 10)
    n_dyn = int((sc.mass > 0).sum())
    t0 = time.time()
    w = World(sc, broadphase="grid")
    sp = w.solver_params()
    sp.m_numIterations = args.iters
    if args.friction is not None:
        sp.m_frictionCoeff = args.friction
    w.set_solver(sp)
    if args.config == "batched":
        w.set_scenes(sc.meta["scene_offsets"])
    setup_s = time.time() - t0
    dt = 1.0 / 60.0
    tuned = None
    for i in range(args.warmup):
        if i == args.warmup // 2 and not args.no_autotune:
            # narrowphase lanes per pair and PGS parts per cell picked by timing
            # trial steps from a saved state image, which is restored afterwards;
            # the rest of the warmup captures the step graph again
            lanes, parts, tms = w.autotune(dt, steps=args.autotune_steps)
            tuned = {"lanes_per_pair": lanes, "pgs_parts": parts, "trial_ms": round(tms, 4)}
        if args.eager:
            w.step(dt)
        else:
            w.step_graph(dt, 1)
    w.sync()
    # timed region: exactly args.steps steps, graph-launched (or eager with --eager)
    w.sync()
    t0 = time.perf_counter()
    if args.eager:
        for _ in range(args.steps):
            w.step(dt)
    else:
        w.step_graph(dt, args.steps)
    w.sync()
    t1 = time.perf_counter()
    sec = t1 - t0
    value = n_dyn * args.steps / sec
    ms_per_step = 1000.0 * sec / args.steps
    # per-stage breakdown from HIP events recorded on the pipeline's stream
    # (eager steps right after the timed region, same state)
    from bullet3_amd.capi import STAGE_NAMES
    w.enable_timing(True)
    reps = 5
    acc = np.zeros(len(STAGE_NAMES))
    for _ in range(reps):
        w.step(dt)
        acc += w.stage_times()
    st = dict(zip(STAGE_NAMES, (acc / reps).tolist()))
    n_pairs = len(w.pairs())
    n_contacts = len(w.contacts())
    launches = {"narrowphase_mpr": 1, "narrowphase_sat": 1, "pgs_normal": 8 * args.iters,
                "pgs_friction": 8 * args.iters}
    kb = kernel_bytes(sc.num_bodies, n_pairs, n_contacts)
    dom = max(launches, key=lambda k: st[k])
    avg_ms = st[dom] / launches[dom]
    achieved = kb[dom] / (avg_ms * 1e-3) / 1e9
    total_bytes = step_bytes(sc.num_bodies, n_pairs, n_contacts, args.iters)
    traffic = args.traffic
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if traffic is None and os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        traffic = t.get("bytes_per_launch", {}).get(dom)
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
        "config": {"workload": workload_name(args, n_dyn),
                   "bodies": n_dyn, "pairs": n_pairs, "contacts": n_contacts, "pgs_iterations": args.iters,
                   "setup_s": round(setup_s, 2), "autotune": tuned},
        "stages_ms": {k: round(float(v), 4) for k, v in st.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "launch_ms": round(avg_ms, 5), "bytes_per_launch": int(kb[dom]),
                     "step_algorithmic_bytes": int(total_bytes),
                     "step_frac": round(total_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
    }
    if not args.no_learned:
        out["warm_start"] = learned_eval(w, args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args)
        if out["cpu_baseline"]["value"] > 0:
            out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    w.close()
    return out


def learned_eval(w, args):
    """PGS iterations to a residual, learned warm start against plain PGS.

    From the pile's post-benchmark state: log every contact point of
    --learn-collect-steps steps of plain previous-frame warm-start PGS at
    --learn-collect-iters iterations (near-converged impulses as targets), train
    the network on the device (--learn-train-steps Adam steps of 8192 rows),
    then for cold start (the reference), previous-frame and learned warm start
    and each iteration count run --learn-eval-steps steps from that state (no
    impulse history at the first) and average the last normal sweep's L2
    residual over the steps after the first --learn-skip.  Reported: the
    iterations each mode needs to reach plain 20-iteration PGS's residual and
    1e-4, and graph-timed ms/step at those counts."""
    from bullet3_amd import warmstart as ws
    from bullet3_amd.learned import WarmStartNet
    dt = 1.0 / 60.0
    state = w.bodies()
    sp0 = w.solver_params()
    t0 = time.time()
    X, T, rows = ws.collect(w, args.learn_collect_steps, args.learn_capacity, dt=dt, mode=ws.PREVIOUS,
                            iterations=args.learn_collect_iters)
    collect_s = time.time() - t0
    net = WarmStartNet(24, 3, max_train_rows=8192, seed=1)
    t0 = time.time()
    hist = ws.train(net, X, T, rows, args.learn_train_steps, batch=8192, lr=1e-3, lr_final=1e-4, normalize=True)
    train_s = time.time() - t0
    X.free()
    T.free()
    iters = [2, 3, 4, 5, 6, 8, 10, 12, 14, 16, 18, 20]
    curves = {}
    for mode in (ws.COLD, ws.PREVIOUS, ws.LEARNED):
        curves[mode] = {}
        for k in iters:
            curves[mode][k], _ = ws.run_residuals(w, state, mode, k, args.learn_eval_steps, args.learn_skip, dt=dt,
                                                  net=net)
    name = ws.MODE_NAMES
    targets = {"plain_cold_20": curves[ws.COLD][20], "plain_warm_20": curves[ws.PREVIOUS][20], "1e-4": 1e-4}
    k_to = {t: {name[m]: ws.iterations_to(curves[m], r) for m in curves} for t, r in targets.items()}
    ms = {}

    def timed(mode, k):
        key = "%s_%d" % (name[mode], k)
        if key not in ms:
            ms[key] = round(ws.time_steps(w, state, mode, k, args.learn_time_steps, dt=dt), 4)
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
    w.set_bodies(state)
    w.clear_warm_history()
    w.set_solver(sp0)
    net.close()
    return {
        "residual": "L2 of the last normal sweep, mean over steps",
        "targets": {t: float("%.4e" % r) for t, r in targets.items()},
        "iterations_to": k_to,
        "learned_vs_plain_20": speed,
        "ms_per_step": ms,
        "curves": {name[m]: {str(k): float("%.4e" % v) for k, v in curves[m].items()} for m in curves},
        "training": {"rows": int(rows), "collect_steps": args.learn_collect_steps,
                     "collect_iterations": args.learn_collect_iters, "collect_s": round(collect_s, 2),
                     "adam_steps": args.learn_train_steps, "batch": 8192, "train_s": round(train_s, 2),
                     "loss_first_100": float(np.mean(hist[:100])) if len(hist) else None,
                     "loss_last_100": float(np.mean(hist[-100:])) if len(hist) else None},
    }


def cpu_baseline(w, args):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import OracleWorld
    ow = OracleWorld(w)
    dt = 1.0 / 60.0
    n_dyn = int((ow.bodies["invMass"] > 0).sum())
    t0 = time.perf_counter()
    ow.step(dt)
    one = time.perf_counter() - t0
    steps = max(1, int(args.cpu_seconds / max(one, 1e-6)))
    t0 = time.perf_counter()
    for _ in range(steps):
        ow.step(dt)
    sec = time.perf_counter() - t0
    return {"value": round(n_dyn * steps / sec, 1), "unit": "body-steps/s", "cores": 1, "kind": "port",
            "sample": "%d oracle steps of the same pile from the GPU's post-warmup state (%.1f s)" % (steps, sec)}


def calibrate_copy(mb):
    import ctypes
    from bullet3_amd import capi
    nbytes = mb << 20
    a, b = capi.DeviceBuffer(nbytes), capi.DeviceBuffer(nbytes)
    capi.call("b3h_memcpy_dtod", ctypes.c_void_p(b.ptr), ctypes.c_void_p(a.ptr), nbytes)
    capi.call("b3h_synchronize")
    print(json.dumps({"calibration_copy_bytes": nbytes}), flush=True)
    a.free()
    b.free()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--friction", type=float, default=None)
    ap.add_argument("--config", default="pile", choices=["pile", "cubes", "batched"])
    ap.add_argument("--scenes", type=int, default=4096, help="scenes of the batched workload")
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=32)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start evaluation")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default narrowphase lanes per pair and PGS parts instead of timing the choices")
    ap.add_argument("--autotune-steps", type=int, default=3)
    ap.add_argument("--eager", action="store_true",
                    help="launch every step's kernels one by one instead of replaying a captured HIP graph")
    ap.add_argument("--learn-collect-steps", type=int, default=40)
    ap.add_argument("--learn-capacity", type=int, default=60_000_000)
    ap.add_argument("--learn-train-steps", type=int, default=12000)
    ap.add_argument("--learn-collect-iters", type=int, default=60)
    ap.add_argument("--learn-eval-steps", type=int, default=16)
    ap.add_argument("--learn-skip", type=int, default=6)
    ap.add_argument("--learn-time-steps", type=int, default=30)
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="after the run, one device-to-device copy of this many MiB (a plain copy kernel of known "
                         "size, to scale rocprof FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per launch of the dominant kernel from rocprof FETCH_SIZE+WRITE_SIZE runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path here (the step would shard by scene "
              "or by spatial slab); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    print(json.dumps(run(args)), flush=True)
    if args.calibrate_copy_mb > 0:
        calibrate_copy(args.calibrate_copy_mb)


if __name__ == "__main__":
    main()
