"""State dump / load through the C ABI (b3h_pipeline_save_state /
b3h_pipeline_load_state): stepping after a load repeats, bit for bit, what
stepping after the save did -- eager and graph-launched, cold start and
previous-frame warm start (whose history travels in the image) -- and bad
images are refused with an error instead of being half applied."""
import ctypes

import numpy as np
import pytest

from bullet3_amd import capi, scenes
from bullet3_amd.world import World
from helpers import bits_equal, first_diff

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


@pytest.fixture(scope="module", autouse=True)
def need_device():
    if capi.device_count() < 1:
        pytest.skip("no HIP device")


def pile():
    return scenes.convex_pile(6, 4, 6, spacing=1.0, n_hull_shapes=8, drop=0.0, seed=7)


def make_world(mode, iters=10):
    w = World(pile(), broadphase="grid")
    sp = w.solver_params()
    sp.m_numIterations = iters
    sp.m_warmStartMode = mode
    w.set_solver(sp)
    return w


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("graph", [False, True])
def test_resume_repeats_bit_for_bit(mode, graph):
    print("mode %d graph %d" % (mode, graph), flush=True)
    w = make_world(mode)
    for _ in range(12):
        w.step(DT)
    image = w.save_state()
    assert len(w.contacts()) > 0

    def run():
        if graph:
            w.step_graph(DT, 8)
        else:
            for _ in range(8):
                w.step(DT)
        w.sync()
        return w.bodies()

    a = run()
    w.load_state(image)
    b = run()
    assert bits_equal(a, b), first_diff(a, b)
    # and into a second pipeline with the same registrations
    w2 = make_world(0, iters=3)
    w2.load_state(image)
    assert w2.solver_params().m_warmStartMode == mode
    assert w2.solver_params().m_numIterations == 10
    c = None
    if graph:
        w2.step_graph(DT, 8)
        w2.sync()
        c = w2.bodies()
    else:
        for _ in range(8):
            w2.step(DT)
        c = w2.bodies()
    assert bits_equal(a, c), first_diff(a, c)
    w2.close()
    w.close()


def test_history_size_in_image():
    w = make_world(1)
    for _ in range(5):
        w.step(DT)
    n = len(w.contacts())
    image = w.save_state()
    header = np.frombuffer(image[:24], np.int32)
    assert bytes(image[:8]) == b"B3HSTATE"
    nb, nh = int(header[4]), int(header[5])
    assert nb == w.num_bodies and nh == n > 0
    w.close()


def test_bad_images_are_refused():
    w = make_world(1)
    for _ in range(3):
        w.step(DT)
    image = w.save_state()
    before = w.bodies()
    lib = capi.lib()

    def load(img):
        buf = np.frombuffer(img, np.uint8).copy()
        return lib.b3h_pipeline_load_state(w.pl, capi.ptr(buf), len(buf))

    assert load(image[:10]) == -1                                # shorter than the header
    assert load(b"X" + image[1:]) == -1                          # magic
    assert load(image[:-4]) == -5                                # truncated payload
    assert load(image + b"\0" * 16) == -5                        # trailing bytes
    small = World(scenes.convex_pile(3, 2, 3, spacing=1.0, n_hull_shapes=8, drop=0.0, seed=7), broadphase="grid")
    buf = np.frombuffer(image, np.uint8).copy()
    assert lib.b3h_pipeline_load_state(small.pl, capi.ptr(buf), len(buf)) == -5   # body count
    small.close()
    assert bits_equal(w.bodies(), before)                        # nothing applied
    n = ctypes.c_size_t()
    tiny = np.zeros(16, np.uint8)
    assert lib.b3h_pipeline_save_state(w.pl, capi.ptr(tiny), 16, ctypes.byref(n)) == -2
    assert n.value == len(image)
    w.close()


def test_autotune_leaves_the_simulation_unchanged():
    """the autotuner times trial steps from a state image and restores it, and
    its knobs do not change results: stepping afterwards equals a twin world
    that never tuned, bit for bit"""
    w1, w2 = make_world(1), make_world(1)
    for _ in range(10):
        w1.step(DT)
        w2.step(DT)
    lanes, parts, ms = w2.autotune(DT, steps=2)
    print("autotune: lanes %d parts %d %.3f ms" % (lanes, parts, ms), flush=True)
    assert lanes in (16, 32, 64) and parts in (1, 4, 8) and ms > 0
    assert bits_equal(w1.bodies(), w2.bodies())
    for _ in range(6):
        w1.step(DT)
        w2.step(DT)
    a, b = w1.bodies(), w2.bodies()
    assert bits_equal(a, b), first_diff(a, b)
    w1.close()
    w2.close()


def test_graph_recaptured_after_component_changes():
    """a captured step is re-captured when a setter changes what it refers to:
    grid parameters (whose scratch an eager step then reallocates), gravity,
    residual tracking, narrowphase lanes per pair; replays keep matching an
    eager twin"""
    a, b = make_world(0), make_world(0)
    cs = np.full(3, a.cell_size * 1.25, np.float32)
    dims = np.array([64, 64, 64], np.int32)
    for rep, change in enumerate(["grid", "gravity", "residual", "lanes"]):
        a.step_graph(DT, 3)
        for _ in range(3):
            b.step(DT)
        a.sync()
        assert bits_equal(a.bodies(), b.bodies()), "before %s" % change
        for w in (a, b):
            if change == "grid":
                capi.call("b3h_bp_set_grid_params", w.bp, capi.ptr(cs), capi.ptr(dims), 256)
            elif change == "gravity":
                capi.call("b3h_pipeline_set_gravity", w.pl, capi.ptr(np.array([0.5, -12.0, 0.0], np.float32)))
            elif change == "residual":
                w.track_residual(True)
            else:
                capi.call("b3h_np_set_pairs_per_group", w.np_, 16)
        a.step(DT)
        b.step(DT)
        a.step_graph(DT, 3)
        for _ in range(3):
            b.step(DT)
        a.sync()
        assert bits_equal(a.bodies(), b.bodies()), "after %s\n%s" % (change, first_diff(a.bodies(), b.bodies()))
    a.close()
    b.close()


def test_autotune_leaves_contact_log_alone():
    w = make_world(1)
    for _ in range(5):
        w.step(DT)
    lx, lt = capi.DeviceBuffer(4 * 24 * 200000), capi.DeviceBuffer(4 * 3 * 200000)
    w.set_contact_log(lx.ptr, lt.ptr, 200000)
    w.step(DT)
    rows = w.contact_log_count()
    assert rows > 0
    w.autotune(DT, steps=2)
    assert w.contact_log_count() == rows
    w.set_contact_log(None, None, 0)
    lx.free()
    lt.free()
    w.close()
