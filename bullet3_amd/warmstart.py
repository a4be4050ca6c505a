"""Learned warm start workflow on the device: log contacts from plain PGS,
train the network with its own forward/backward/Adam kernels, and measure how
many PGS iterations each initial iterate needs to reach a residual.

Modes (b3hSolverParams::m_warmStartMode): 0 cold start, as the reference
(b3ConvertConstraint4.h:73,107,132 zero the impulses); 1 previous-frame
impulses; 2 previous-frame impulses corrected by the network.  Only the
initial iterate differs, so the solve's fixed point does not.

Residual of a sweep: sqrt(sum of squared impulse changes / sum of squared
impulses) over the rows it updates; a step's residual is that of its last
normal sweep.
"""
import time

import numpy as np

from .learned import DeviceArray, WarmStartNet

COLD, PREVIOUS, LEARNED = 0, 1, 2
MODE_NAMES = {COLD: "cold", PREVIOUS: "previous_frame", LEARNED: "learned"}


def collect(world, steps, capacity_rows, dt=1.0 / 60.0, mode=COLD, iterations=20):
    """run `steps` steps with the given solver mode and log every contact point;
    returns (features DeviceArray [cap][24], targets DeviceArray [cap][3], rows)"""
    X = DeviceArray((capacity_rows, 24))
    T = DeviceArray((capacity_rows, 3))
    sp = world.solver_params()
    sp.m_numIterations = iterations
    sp.m_warmStartMode = mode
    world.set_solver(sp)
    world.set_contact_log(X.ptr, T.ptr, capacity_rows)
    for _ in range(steps):
        world.step(dt)
    rows = world.contact_log_count()
    world.set_contact_log(None, None, 0)
    return X, T, rows


def collect_into(world, X, T, row0, steps, dt=1.0 / 60.0, mode=PREVIOUS, iterations=20):
    """like collect, appending to existing logs X [cap][24] / T [cap][3] from
    row `row0`; returns the new row count"""
    cap = X.shape[0] - row0
    if cap <= 0 or steps <= 0:
        return row0
    sp = world.solver_params()
    sp.m_numIterations = iterations
    sp.m_warmStartMode = mode
    world.set_solver(sp)
    world.set_contact_log(X.ptr + 4 * 24 * row0, T.ptr + 4 * 3 * row0, cap)
    for _ in range(steps):
        world.step(dt)
    rows = world.contact_log_count()
    world.set_contact_log(None, None, 0)
    return row0 + rows


def fit_input_normalization(net, X, rows, sample=400_000):
    """standardise every feature: shift = mean, scale = 1 / std over the first
    `sample` logged rows (constant features keep scale 1)"""
    x = X.numpy(min(rows, sample)).astype(np.float64)
    mean = x.mean(0)
    std = x.std(0)
    scale = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 1.0)
    net.set_input_normalization(mean.astype(np.float32), scale.astype(np.float32))
    return mean, scale


def train(net, X, T, rows, steps, batch=8192, lr=1e-3, lr_final=1e-4, seed=1, chunk=250, normalize=True):
    """Adam on minibatches drawn from the log, learning rate decayed
    geometrically from lr to lr_final; returns the loss of every step.  With
    `normalize`, the network's input normalisation is fitted to the log first."""
    if normalize and rows > 0:
        fit_input_normalization(net, X, rows)
    hist = []
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        frac = done / max(1, steps - 1)
        rate = lr * (lr_final / lr) ** frac
        hist.append(net.train_dev(X.ptr, 24, T.ptr, 3, rows, batch, k, rate, seed + done))
        done += k
    return np.concatenate(hist) if hist else np.zeros(0, np.float32)


def run_residuals(world, state, mode, iterations, steps, skip, dt=1.0 / 60.0, net=None):
    """from `state` (bodies) with no impulse history, `steps` steps at the given
    mode and iteration count; mean and per-step residuals of steps >= skip"""
    world.set_bodies(state)
    world.clear_warm_history()
    if mode == LEARNED and net is not None:
        world.set_warm_start(LEARNED, net)
    sp = world.solver_params()
    sp.m_numIterations = iterations
    sp.m_warmStartMode = mode
    world.set_solver(sp)
    world.track_residual(True)
    r = []
    for s in range(steps):
        world.step(dt)
        if s >= skip:
            r.append(float(world.residuals_l2()[iterations - 1]))
    world.track_residual(False)
    return float(np.mean(r)), r


def time_steps(world, state, mode, iterations, steps, dt=1.0 / 60.0, warmup=5):
    """ms per graph-launched step at the given mode and iteration count"""
    world.set_bodies(state)
    world.clear_warm_history()
    sp = world.solver_params()
    sp.m_numIterations = iterations
    sp.m_warmStartMode = mode
    world.set_solver(sp)
    world.step_graph(dt, warmup)
    world.sync()
    t0 = time.perf_counter()
    world.step_graph(dt, steps)
    world.sync()
    return 1000.0 * (time.perf_counter() - t0) / steps


def iterations_to(curve, target):
    """smallest iteration count whose mean residual is <= target (None if none)"""
    ok = [k for k, v in sorted(curve.items()) if v <= target]
    return ok[0] if ok else None
