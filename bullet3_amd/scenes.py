"""Seeded scene generators for the benchmark configurations.

A scene is plain numpy data: hull point clouds (`shapes`), and per body a mass,
position, orientation (x,y,z,w) and shape index.  Body 0 is the static ground
box, registered first as in the reference demos
(examples/OpenCL/rigidbody/GpuConvexScene.cpp:280-304), so the solver's
static0Index is 0.
"""
import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Scene:
    name: str
    shapes: list            # list of (n,3) float32 point clouds
    mass: np.ndarray        # (N,) float32, 0 = static
    pos: np.ndarray         # (N,3) float32
    orn: np.ndarray         # (N,4) float32, x y z w
    shape: np.ndarray       # (N,) int32 index into shapes
    cell_size: float = 3.0  # uniform-grid cell (reference default 3 m)
    meta: dict = field(default_factory=dict)

    @property
    def num_bodies(self):
        return len(self.mass)


def box_points(hx, hy, hz):
    return np.array([[sx * hx, sy * hy, sz * hz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float32)


def sphere_points(rng, n, radius=0.5):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return (v * radius).astype(np.float32)


def quat_from_euler_small(rng, n, max_deg):
    """rotations of up to +-max_deg about each axis (x, then y, then z)"""
    a = np.radians(rng.uniform(-max_deg, max_deg, size=(n, 3)))
    out = np.empty((n, 4), np.float64)
    for i in range(n):
        q = np.array([0, 0, 0, 1.0])
        for ax in range(3):
            h = a[i, ax] * 0.5
            r = np.zeros(4)
            r[ax] = math.sin(h)
            r[3] = math.cos(h)
            q = quat_mul(r, q)
        out[i] = q
    return out.astype(np.float32)


def quat_mul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + bw * ax + ay * bz - az * by,
                     aw * by + bw * ay + az * bx - ax * bz,
                     aw * bz + bw * az + ax * by - ay * bx,
                     aw * bw - ax * bx - ay * by - az * bz])


def random_quats(rng, n):
    """uniform random rotations (Shoemake)"""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    return q.astype(np.float32)


def cube_grid(nx=10, ny=10, nz=10, spacing=1.05, half=0.5, drop=2.0, jitter=0.02, rot_deg=2.0,
              ground_half=(100.0, 0.5, 100.0), seed=1):
    """Reference CPU scene: a grid of unit cubes dropped onto a static box.

    Default = 10x10x10 cubes of half-extent 0.5, mass 1, spacing 1.05 m, seeded
    jitter +-0.02 m and +-2 deg, lowest layer `drop` m above a 200x1x200 m ground."""
    rng = np.random.default_rng(seed)
    shapes = [box_points(*ground_half), box_points(half, half, half)]
    n = nx * ny * nz
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    x = (ii - (nx - 1) / 2.0) * spacing
    z = (kk - (nz - 1) / 2.0) * spacing
    y = drop + half + jj * spacing
    pos = np.stack([x, y, z], 1) + rng.uniform(-jitter, jitter, size=(n, 3))
    orn = quat_from_euler_small(rng, n, rot_deg)
    mass = np.ones(n, np.float32)
    return _with_ground("cube_grid_%dx%dx%d" % (nx, ny, nz), shapes, mass, pos, orn, np.ones(n, np.int32),
                        ground_half, cell_size=3.0)


def _with_ground(name, shapes, mass, pos, orn, shape, ground_half, cell_size, meta=None):
    gpos = np.array([[0.0, -ground_half[1], 0.0]], np.float32)
    return Scene(name=name, shapes=shapes,
                 mass=np.concatenate([[0.0], mass]).astype(np.float32),
                 pos=np.concatenate([gpos, pos]).astype(np.float32),
                 orn=np.concatenate([[[0, 0, 0, 1.0]], orn]).astype(np.float32),
                 shape=np.concatenate([[0], shape]).astype(np.int32),
                 cell_size=cell_size, meta=meta or {})


def convex_pile(nx=64, ny=32, nz=64, spacing=1.2, n_hull_shapes=64, drop=1.0, seed=2, ground_half=(200.0, 0.5, 200.0)):
    """Single large pile: nx*ny*nz dynamic bodies (default 131072), one third
    each unit boxes, 16-point and 32-point random hulls (points uniform on a
    sphere of radius 0.5, a pool of `n_hull_shapes` instanced shapes), masses
    log-uniform in [0.5, 4] kg, lattice spacing 1.2 m, random orientations,
    grid cell = 2 x the largest world-AABB extent (sqrt(3) for these shapes)."""
    rng = np.random.default_rng(seed)
    shapes = [box_points(*ground_half), box_points(0.5, 0.5, 0.5)]
    half = n_hull_shapes // 2
    for k in range(n_hull_shapes):
        shapes.append(sphere_points(rng, 16 if k < half else 32))
    n = nx * ny * nz
    kind = np.arange(n) % 3
    rng.shuffle(kind)
    shape = np.where(kind == 0, 1, np.where(kind == 1, 2 + rng.integers(0, half, n), 2 + half + rng.integers(0, half, n)))
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    x = (ii - (nx - 1) / 2.0) * spacing
    z = (kk - (nz - 1) / 2.0) * spacing
    y = drop + 0.5 + jj * spacing
    pos = np.stack([x, y, z], 1) + rng.uniform(-0.02, 0.02, size=(n, 3))
    orn = random_quats(rng, n)
    mass = np.exp(rng.uniform(np.log(0.5), np.log(4.0), n)).astype(np.float32)
    return _with_ground("convex_pile_%d" % n, shapes, mass, pos, orn, shape.astype(np.int32), ground_half,
                        cell_size=2.0 * math.sqrt(3.0))


def compound_scene(n=64, seed=7, plane=True, sphere=True, ground_half=(12.0, 0.5, 12.0)):
    """Compounds among hulls: a ground box (body 0), optionally a static plane
    at y = 0 and spheres, and n dynamic bodies cycling through a unit box, a
    16-point hull and three compounds: a dumbbell (two small boxes), a three-
    part cluster (a hull, a small box and a unit box under child rotations) and
    a one-child compound whose child sits at the body origin.  Shapes are
    {"compound": [(shape k, position, orientation), ...]} entries after the
    hulls they use."""
    rng = np.random.default_rng(seed)
    qa, qb, qc = (random_quats(rng, 1)[0].astype(np.float32) for _ in range(3))
    shapes = [box_points(*ground_half), box_points(0.5, 0.5, 0.5), sphere_points(rng, 16),
              box_points(0.3, 0.15, 0.15),
              {"compound": [(3, (-0.45, 0.0, 0.0), (0, 0, 0, 1)), (3, (0.45, 0.0, 0.0), (0, 0, 0, 1))]},
              {"compound": [(2, (0.0, 0.35, 0.0), tuple(qa)), (3, (0.4, -0.2, 0.0), tuple(qb)),
                            (1, (-0.4, -0.2, 0.1), tuple(qc))]},
              {"compound": [(1, (0.0, 0.0, 0.0), (0, 0, 0, 1))]}]
    kinds = [1, 2, 4, 5, 6]
    plane_shape = -1
    if plane:
        plane_shape = len(shapes)
        shapes.append({"plane": (0.0, 1.0, 0.0, 0.0)})
    if sphere:
        shapes.append({"sphere": 0.35})
        kinds.append(len(shapes) - 1)
    mass, pos, shape = [], [], []
    side = max(2, int(math.ceil(n ** (1.0 / 3.0))))
    for k in range(n):
        x, z, y = k % side, (k // side) % side, k // (side * side)
        pos.append((1.4 * (x - (side - 1) / 2.0) + rng.uniform(-0.05, 0.05), 0.9 + 1.3 * y,
                    1.4 * (z - (side - 1) / 2.0) + rng.uniform(-0.05, 0.05)))
        shape.append(kinds[k % len(kinds)])
        mass.append(float(np.exp(rng.uniform(np.log(0.5), np.log(2.0)))))
    orn = random_quats(rng, n).astype(np.float32)
    if plane:
        # the plane as the second static body
        pos.insert(0, (0.0, 0.0, 0.0))
        shape.insert(0, plane_shape)
        mass.insert(0, 0.0)
        orn = np.concatenate([[[0, 0, 0, 1.0]], orn]).astype(np.float32)
    return _with_ground("compounds", shapes, np.array(mass, np.float32), np.array(pos, np.float32), orn,
                        np.array(shape, np.int32), ground_half, cell_size=4.5)


def lattice_joints(n, count, first=2, spacing=1.4):
    """point-to-point joints between x-neighbours of compound_scene's lattice
    (dynamic body k is body first + k; neighbours sit spacing apart in x):
    up to count (a, b, pivot in a, pivot in b) tuples, pivots at the midpoint"""
    side = max(2, int(math.ceil(n ** (1.0 / 3.0))))
    out = []
    h = 0.5 * spacing
    for k in range(n - 1):
        if len(out) >= count:
            break
        if k % side != side - 1:
            out.append((first + k, first + k + 1, (h, 0.0, 0.0), (-h, 0.0, 0.0)))
    return out


def aabb_cloud(n=1 << 20, mean_overlap=8.0, hmin=0.05, hmax=1.0, seed=3, side=None):
    """Broadphase stress: n cubic AABBs, half-extents log-uniform [hmin,hmax],
    centres uniform in a cube whose side gives about `mean_overlap` partners
    per box: L^3 = n E[(h1+h2)^3] * 8 / mean_overlap."""
    rng = np.random.default_rng(seed)
    h = np.exp(rng.uniform(np.log(hmin), np.log(hmax), n))
    if side is None:
        lo, hi = np.log(hmin), np.log(hmax)

        def moment(k):
            return (hmax ** k - hmin ** k) / (k * (hi - lo))

        e3 = 2 * moment(3) + 6 * moment(2) * moment(1)
        side = (n * e3 * 8.0 / mean_overlap) ** (1.0 / 3.0)
    c = rng.uniform(-side / 2, side / 2, size=(n, 3))
    mn = (c - h[:, None]).astype(np.float32)
    mx = (c + h[:, None]).astype(np.float32)
    return mn, mx, float(side)


def batched_scenes(n_scenes=4096, min_bodies=16, max_bodies=512, spacing=27.0, n_hull_shapes=32, seed=4):
    """Many independent scenes packed into one world: scene k sits on its own
    static ground (16 x 1 x 16 m) at a lattice position `spacing` m from its
    neighbours, so no two scenes interact.  The solver's 256 batching cells
    (6 m, wrapping every 48 m in x and z) serialise the contacts of one cell;
    at a 27 m pitch the scene centres fall on 16 different offsets within the
    48 m period and spread over all cells, where a pitch dividing 48 m (24 m)
    would put every scene's centre on the same few cells.  Body counts are log-uniform in
    [min_bodies, max_bodies]; even scenes are piles (boxes and 8-24-point hulls,
    random orientations, lattice spacing 1.2 m), odd scenes are box stacks
    (columns, spacing 1.02 m).  Returns one Scene; meta["scene_of_body"] maps
    every body to its scene, meta["scene_offsets"] gives each scene's body range
    (its ground first, then its bodies)."""
    rng = np.random.default_rng(seed)
    shapes = [box_points(8.0, 0.5, 8.0), box_points(0.5, 0.5, 0.5)]
    for k in range(n_hull_shapes):
        shapes.append(sphere_points(rng, int(rng.integers(8, 25))))
    side = int(math.ceil(math.sqrt(n_scenes)))
    counts = np.exp(rng.uniform(math.log(min_bodies), math.log(max_bodies), n_scenes)).round().astype(int)
    mass, pos, orn, shape, owner = [], [], [], [], []
    for k in range(n_scenes):
        ox = (k % side - (side - 1) / 2.0) * spacing
        oz = (k // side - (side - 1) / 2.0) * spacing
        mass.append(0.0)
        pos.append((ox, -0.5, oz))
        orn.append((0.0, 0.0, 0.0, 1.0))
        shape.append(0)
        owner.append(k)
        n = int(counts[k])
        if k % 2 == 0:
            w = int(min(8, math.ceil(math.sqrt(n))))
            q = random_quats(rng, n)
            for i in range(n):
                x, z, y = i % w, (i // w) % w, i // (w * w)
                pos.append((ox + (x - (w - 1) / 2.0) * 1.2, 1.0 + y * 1.2, oz + (z - (w - 1) / 2.0) * 1.2))
                orn.append(tuple(q[i]))
                shape.append(1 if rng.random() < 0.4 else 2 + int(rng.integers(0, n_hull_shapes)))
                mass.append(float(np.exp(rng.uniform(math.log(0.5), math.log(4.0)))))
                owner.append(k)
        else:
            w = int(min(6, math.ceil(math.sqrt(n / 8.0))))
            for i in range(n):
                x, z, y = i % w, (i // w) % w, i // (w * w)
                pos.append((ox + (x - (w - 1) / 2.0) * 1.2, 0.5 + y * 1.02, oz + (z - (w - 1) / 2.0) * 1.2))
                orn.append((0.0, 0.0, 0.0, 1.0))
                shape.append(1)
                mass.append(1.0)
                owner.append(k)
    return Scene(name="batched_%d" % n_scenes, shapes=shapes, mass=np.array(mass, np.float32),
                 pos=np.array(pos, np.float32), orn=np.array(orn, np.float32), shape=np.array(shape, np.int32),
                 cell_size=2.0 * math.sqrt(3.0), meta={"scene_of_body": np.array(owner, np.int32),
                                                       "bodies_per_scene": counts,
                                                       "scene_offsets": np.concatenate(
                                                           [[0], np.cumsum(counts + 1)]).astype(np.int32)})
